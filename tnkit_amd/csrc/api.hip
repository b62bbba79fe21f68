// Library-wide entries: version and per-thread error text.
#include <cstdarg>

#include "tnh_common.h"

namespace tnh {

namespace {
thread_local char g_err[512] = "";
}

void set_error(const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

}  // namespace tnh

extern "C" {

int tnh_version(void) { return 100; }

const char* tnh_last_error(void) { return tnh::g_err; }

// hipGraph capture of a sequence of library calls on one stream (the C-side counterpart of the
// torch-level graph helpers): begin, issue calls, end -> an executable graph to launch and destroy.
int tnh_graph_begin(tnh_stream_t stream) {
  TNH_REQUIRE(stream != nullptr, TNH_ERR_ARG, "tnh_graph_begin: the null stream cannot be captured");
  const hipError_t e =
      hipStreamBeginCapture(static_cast<hipStream_t>(stream), hipStreamCaptureModeThreadLocal);
  TNH_REQUIRE(e == hipSuccess, TNH_ERR_HIP, "tnh_graph_begin: %s", hipGetErrorString(e));
  return TNH_OK;
}

int tnh_graph_end(tnh_stream_t stream, void** graph) {
  TNH_REQUIRE(stream != nullptr && graph != nullptr, TNH_ERR_ARG, "tnh_graph_end: null argument");
  *graph = nullptr;
  hipGraph_t g = nullptr;
  hipError_t e = hipStreamEndCapture(static_cast<hipStream_t>(stream), &g);
  TNH_REQUIRE(e == hipSuccess && g != nullptr, TNH_ERR_HIP, "tnh_graph_end: capture: %s",
              hipGetErrorString(e));
  hipGraphExec_t x = nullptr;
  e = hipGraphInstantiate(&x, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  TNH_REQUIRE(e == hipSuccess, TNH_ERR_HIP, "tnh_graph_end: instantiate: %s",
              hipGetErrorString(e));
  *graph = x;
  return TNH_OK;
}

int tnh_graph_launch(void* graph, tnh_stream_t stream) {
  TNH_REQUIRE(graph != nullptr, TNH_ERR_ARG, "tnh_graph_launch: null graph");
  const hipError_t e = hipGraphLaunch(static_cast<hipGraphExec_t>(graph),
                                      static_cast<hipStream_t>(stream));
  TNH_REQUIRE(e == hipSuccess, TNH_ERR_HIP, "tnh_graph_launch: %s", hipGetErrorString(e));
  return TNH_OK;
}

int tnh_graph_destroy(void* graph) {
  if (graph == nullptr) return TNH_OK;
  const hipError_t e = hipGraphExecDestroy(static_cast<hipGraphExec_t>(graph));
  TNH_REQUIRE(e == hipSuccess, TNH_ERR_HIP, "tnh_graph_destroy: %s", hipGetErrorString(e));
  return TNH_OK;
}

}  // extern "C"
