// The two-site matvec for callers that cannot allocate device memory (code under HIP graph
// capture, tight serving loops): tnh_apply_two_site_fused with its workspace taken from a
// file-scope device array instead of a caller buffer.
//
// The array lives in this translation unit only, so only a process that calls this entry loads
// it (TNH_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends). Each GPU holds its own
// copy at its own address; the address is looked up once per device with hipGetSymbolAddress
// and kept. Every call on one GPU shares that GPU's array: calls that overlap on one GPU (two
// streams, or two host threads) race on it. The matvec's kernels write every workspace element
// they read, so nothing is cleared between calls.
#include <mutex>

#include "tnh_common.h"

namespace {

static __device__ __align__(256) unsigned char scratch[TNH_NOALLOC_SCRATCH_BYTES];

constexpr int kMaxDevices = 64;
void* g_scratch[kMaxDevices] = {};
std::mutex g_scratch_mu;

// The device address of `scratch` on the current device (looked up on the first call there).
int scratch_address(void** out) {
  int dev = -1;
  hipError_t e = hipGetDevice(&dev);
  TNH_REQUIRE(e == hipSuccess, TNH_ERR_HIP, "tnh_apply_two_site_fused_noalloc: hipGetDevice: %s",
              hipGetErrorString(e));
  TNH_REQUIRE(dev >= 0 && dev < kMaxDevices, TNH_ERR_UNSUPPORTED,
              "tnh_apply_two_site_fused_noalloc: device %d beyond the %d supported", dev,
              kMaxDevices);
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (g_scratch[dev] == nullptr) {
    void* p = nullptr;
    e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    TNH_REQUIRE(e == hipSuccess && p != nullptr, TNH_ERR_HIP,
                "tnh_apply_two_site_fused_noalloc: hipGetSymbolAddress(scratch) on device %d: %s",
                dev, hipGetErrorString(e));
    g_scratch[dev] = p;
  }
  *out = g_scratch[dev];
  return TNH_OK;
}

}  // namespace

extern "C" int tnh_apply_two_site_fused_noalloc(int dtype, int io_dtype, int a, int d1, int d2,
                                                int b, int wl, int wr, const void* LT,
                                                const void* S, const void* RT, const void* theta,
                                                void* out, int tile1, int tile3,
                                                tnh_stream_t stream) {
  TNH_REQUIRE(a > 0 && b > 0 && wl > 0 && wr > 0 && d1 > 0 && d2 > 0, TNH_ERR_ARG,
              "tnh_apply_two_site_fused_noalloc: sizes must be positive");
  const size_t need = tnh_two_site_ws_bytes(dtype, a, d1, d2, b, wl, wr);
  TNH_REQUIRE(need <= (size_t)TNH_NOALLOC_SCRATCH_BYTES, TNH_ERR_WS,
              "tnh_apply_two_site_fused_noalloc: a=%d d1=%d d2=%d b=%d wl=%d wr=%d needs %zu "
              "bytes of scratch, above the fixed %zu; use tnh_apply_two_site_fused with a "
              "workspace",
              a, d1, d2, b, wl, wr, need, (size_t)TNH_NOALLOC_SCRATCH_BYTES);
  void* ws = nullptr;
  const int rc = scratch_address(&ws);
  if (rc != TNH_OK) return rc;
  return tnh_apply_two_site_fused(dtype, io_dtype, a, d1, d2, b, wl, wr, LT, S, RT, theta, out,
                                  ws, TNH_NOALLOC_SCRATCH_BYTES, tile1, tile3, stream);
}
