// Driver of libtnkit_hip through its C ABI alone (no Python, no torch): times the two-site matvec
// of the XXZ bulk problem, or a plain device copy for counter calibration. Used for rocprofv3
// counter runs and as the example of a C caller of include/tnkit_hip.h.
//
//   cabi_bench mixed|fp64|bf16 [chi=1024] [reps=10] H_eff theta, D = 5, d = 2, random L, R, theta
//   cabi_bench copy [MB=256] [reps=3]               y = 1.0 * x over MB megabytes (tnh_scale_by)
#include <hip/hip_runtime.h>

#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "../../include/tnkit_hip.h"

#define HIP_OK(x)                                                              \
  do {                                                                         \
    hipError_t e_ = (x);                                                       \
    if (e_ != hipSuccess) {                                                    \
      fprintf(stderr, "%s:%d %s\n", __FILE__, __LINE__, hipGetErrorString(e_)); \
      return 2;                                                                \
    }                                                                          \
  } while (0)
#define TNH_OK_(x)                                                         \
  do {                                                                     \
    int rc_ = (x);                                                         \
    if (rc_ != TNH_OK) {                                                   \
      fprintf(stderr, "%s:%d tnh status %d: %s\n", __FILE__, __LINE__, rc_, \
              tnh_last_error());                                           \
      return 3;                                                            \
    }                                                                      \
  } while (0)

static uint64_t g_state = 0x9E3779B97F4A7C15ull;
static double uniform_pm1() {  // xorshift64*, in [-1, 1)
  g_state ^= g_state >> 12;
  g_state ^= g_state << 25;
  g_state ^= g_state >> 27;
  const uint64_t r = g_state * 2685821657736338717ull;
  return (double)(r >> 11) * (2.0 / 9007199254740992.0) - 1.0;
}

static void* upload(const std::vector<double>& h) {
  void* d = nullptr;
  if (hipMalloc(&d, h.size() * sizeof(double)) != hipSuccess) return nullptr;
  if (hipMemcpy(d, h.data(), h.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
    return nullptr;
  return d;
}

int main(int argc, char** argv) {
  const char* mode = argc > 1 ? argv[1] : "mixed";
  hipStream_t st;
  HIP_OK(hipStreamCreate(&st));
  hipEvent_t e0, e1;
  HIP_OK(hipEventCreate(&e0));
  HIP_OK(hipEventCreate(&e1));

  if (strcmp(mode, "copy") == 0) {
    const long mb = argc > 2 ? atol(argv[2]) : 256;
    const int reps = argc > 3 ? atoi(argv[3]) : 3;
    const long n = mb * (1L << 20) / 8;
    void *x, *y;
    HIP_OK(hipMalloc(&x, n * 8));
    HIP_OK(hipMalloc(&y, n * 8));
    HIP_OK(hipMemsetAsync(x, 0, n * 8, st));
    HIP_OK(hipEventRecord(e0, st));
    for (int r = 0; r < reps; ++r)
      TNH_OK_(tnh_scale_by(n, (const double*)x, nullptr, 1.0, (double*)y, st));
    HIP_OK(hipEventRecord(e1, st));
    HIP_OK(hipEventSynchronize(e1));
    float ms = 0;
    HIP_OK(hipEventElapsedTime(&ms, e0, e1));
    printf("{\"mode\": \"copy\", \"bytes_read\": %ld, \"bytes_written\": %ld, \"ms\": %.4f, "
           "\"GBps\": %.1f}\n", n * 8, n * 8, ms / reps, 2.0 * n * 8 / (ms / reps * 1e-3) / 1e9);
    return 0;
  }

  const bool mixed = strcmp(mode, "mixed") == 0;
  const bool bf16 = strcmp(mode, "bf16") == 0;
  if (!mixed && !bf16 && strcmp(mode, "fp64") != 0) {
    fprintf(stderr, "usage: cabi_bench mixed|fp64|bf16 [chi] [reps] | copy [MB] [reps]\n");
    return 1;
  }
  const int chi = argc > 2 ? atoi(argv[2]) : 1024;
  const int reps = argc > 3 ? atoi(argv[3]) : 10;
  const int D = 5, d = 2;
  // XXZ bulk MPO (Jxy = Jz = 1), W[wl][wr][s][s'] with S = sigma / 2
  std::vector<double> W(D * D * d * d, 0.0);
  auto w = [&](int a, int b, int s, int t) -> double& { return W[((a * D + b) * d + s) * d + t]; };
  w(0, 0, 0, 0) = w(0, 0, 1, 1) = 1.0;  // I
  w(0, 1, 0, 1) = 1.0;                  // S+
  w(0, 2, 1, 0) = 1.0;                  // S-
  w(0, 3, 0, 0) = 0.5, w(0, 3, 1, 1) = -0.5;
  w(1, 4, 1, 0) = 0.5;                  // S-/2
  w(2, 4, 0, 1) = 0.5;                  // S+/2
  w(3, 4, 0, 0) = 0.5, w(3, 4, 1, 1) = -0.5;
  w(4, 4, 0, 0) = w(4, 4, 1, 1) = 1.0;
  const long nL = (long)chi * D * chi, nth = (long)chi * d * d * chi;
  std::vector<double> hL(nL), hR(nL), ht(nth);
  for (auto& v : hL) v = uniform_pm1();
  for (auto& v : hR) v = uniform_pm1();
  for (auto& v : ht) v = uniform_pm1();
  void* L = upload(hL);
  void* R = upload(hR);
  void* Wd = upload(W);
  void* th = upload(ht);
  if (!L || !R || !Wd || !th) {
    fprintf(stderr, "device allocation failed\n");
    return 2;
  }
  const int dt = mixed ? TNH_F32 : (bf16 ? TNH_BF16 : TNH_F64);
  const size_t e = mixed ? 4 : (bf16 ? 2 : 8);
  void *Lo = L, *Ro = R, *Wo = Wd;
  if (bf16) {  // L, R in bf16, W (and the built S) in float32
    HIP_OK(hipMalloc(&Lo, nL * 2));
    HIP_OK(hipMalloc(&Ro, nL * 2));
    HIP_OK(hipMalloc(&Wo, W.size() * 4));
    TNH_OK_(tnh_convert(TNH_BF16, TNH_F64, nL, L, Lo, st));
    TNH_OK_(tnh_convert(TNH_BF16, TNH_F64, nL, R, Ro, st));
    TNH_OK_(tnh_convert(TNH_F32, TNH_F64, (long)W.size(), Wd, Wo, st));
  }
  if (mixed) {
    HIP_OK(hipMalloc(&Lo, nL * 4));
    HIP_OK(hipMalloc(&Ro, nL * 4));
    HIP_OK(hipMalloc(&Wo, W.size() * 4));
    TNH_OK_(tnh_convert(TNH_F32, TNH_F64, nL, L, Lo, st));
    TNH_OK_(tnh_convert(TNH_F32, TNH_F64, nL, R, Ro, st));
    TNH_OK_(tnh_convert(TNH_F32, TNH_F64, (long)W.size(), Wd, Wo, st));
  }
  void *LT, *S, *RT, *out, *ws;
  HIP_OK(hipMalloc(&LT, nL * e));
  HIP_OK(hipMalloc(&RT, nL * e));
  HIP_OK(hipMalloc(&S, (size_t)d * d * D * D * d * d * (bf16 ? 4 : e)));
  HIP_OK(hipMalloc(&out, nth * 8));
  const size_t wsb = tnh_two_site_ws_bytes(dt, chi, d, d, chi, D, D);
  HIP_OK(hipMalloc(&ws, wsb));
  TNH_OK_(tnh_fused_two_site(dt, chi, D, D, D, d, d, chi, Lo, Wo, Wo, Ro, LT, S, RT, st));
  const int io = TNH_F64;
  TNH_OK_(tnh_apply_two_site_fused(dt, io, chi, d, d, chi, D, D, LT, S, RT, th, out, ws, wsb, 0, 0,
                                   st));  // warm-up
  HIP_OK(hipEventRecord(e0, st));
  for (int r = 0; r < reps; ++r)
    TNH_OK_(tnh_apply_two_site_fused(dt, io, chi, d, d, chi, D, D, LT, S, RT, th, out, ws, wsb, 0,
                                     0, st));
  HIP_OK(hipEventRecord(e1, st));
  HIP_OK(hipEventSynchronize(e1));
  float ms = 0;
  HIP_OK(hipEventElapsedTime(&ms, e0, e1));
  const double c = chi;
  const double flops = 4.0 * d * d * D * c * c * c + 2.0 * d * d * d * d * D * D * c * c;
  printf("{\"mode\": \"%s\", \"chi\": %d, \"reps\": %d, \"ms\": %.4f, \"TFLOPs\": %.2f}\n", mode,
         chi, reps, ms / reps, flops / (ms / reps * 1e-3) / 1e12);
  return 0;
}
