// Complex128 Krylov-exponential kernels, batched over independent problems (legs): one step of
// tnkit's fixed-m Lanczos exponential (tnkit/eigensolver.py:95-179; batched form
// tnkit/batch_tdvp.py:247-324) with every scalar kept on the device:
//   w -= beta[j-1] V[j-1]                       (j > 0, beta from the previous step)
//   c = V[0..j]^H w;  alpha[j] = Re c[j];  w -= c V      (CGS pass 1: alpha read from it)
//   c = V[0..j]^H w;  w -= c V                            (CGS pass 2)
//   beta[j] = |w|;  V[j+1] = w / max(beta[j], 1e-200)     (guard_denom, tnkit/backend.py:699-708)
// and the final combination out = sum_k c_k V[k]. Reductions are two-stage and deterministic.
#include "cplx.h"
#include "tnh_common.h"

namespace tnh {

namespace {

constexpr int kKMax = 32;      // Krylov dimension held in registers (tnkit's default m is 30)
constexpr int kPartRows = 1024;  // partial-sum rows of the workspace (blocks over all legs)

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Ragged batches: `live` (2 ints per leg, may be null) = the leg's live extents (x, y) in a
// vector laid out as rows of `rowlen` entries whose last index runs over `bpad`: entry i is live
// when i / rowlen < x and i % bpad < y. The others are zero in every vector of the leg, so a pass
// skips their loads and stores (the sums are the same).
struct Ragged {
  const int* live;
  int rowlen, bpad;
};

__device__ __forceinline__ bool live_entry(const Ragged& rg, int lx, int ly, long i) {
  if (rg.live == nullptr) return true;
  const int ii = static_cast<int>(i);
  return ii / rg.rowlen < lx && ii % rg.bpad < ly;
}

// One pass over the live basis of every leg: optional three-term subtraction, subtraction of
// the previous pass's projection (complex coefficients `coef`), then partial sums of the next
// projection <V[k], w> (mode 0) or of |w|^2 (mode 1). part is [leg][block][2*NV] doubles.
template <int NV>
__global__ void __launch_bounds__(256)
    cgs_c_kernel(long n, int nv, const c128* __restrict__ V, long ldv, long sV, c128* __restrict__ w,
                 long sw, const double* __restrict__ coef, long scoef, const double* __restrict__ beta,
                 long sbeta, int beta_row, int mode, double* __restrict__ part, int nbl, Ragged rg) {
  const int leg = blockIdx.y;
  V += leg * sV;
  w += leg * sw;
  const int lx = rg.live ? rg.live[2 * leg] : 0, ly = rg.live ? rg.live[2 * leg + 1] : 0;
  __shared__ double cs[2 * NV];
  __shared__ double red[4][2 * NV];
  for (int k = threadIdx.x; k < 2 * NV; k += 256)
    cs[k] = (coef && k < 2 * nv) ? coef[leg * scoef + k] : 0.0;
  __syncthreads();
  const double bsub = beta ? beta[leg * sbeta] : 0.0;
  const bool writes = (coef != nullptr) || (beta != nullptr);
  double ar[NV], ai[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) ar[k] = ai[k] = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)nbl * 256) {
    if (!live_entry(rg, lx, ly, i)) continue;
    c128 wi = w[i];
    if (beta) {
      const c128 vp = V[(long)beta_row * ldv + i];
      wi.re -= bsub * vp.re;
      wi.im -= bsub * vp.im;
    }
    c128 v[NV];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      v[k] = (k < nv) ? V[(long)k * ldv + i] : c128();
      const double cr = cs[2 * k], ci = cs[2 * k + 1];
      wi.re -= cr * v[k].re - ci * v[k].im;
      wi.im -= cr * v[k].im + ci * v[k].re;
    }
    if (writes) w[i] = wi;
    if (mode == 0) {
#pragma unroll
      for (int k = 0; k < NV; ++k) {  // conj(v) * w
        ar[k] += v[k].re * wi.re + v[k].im * wi.im;
        ai[k] += v[k].re * wi.im - v[k].im * wi.re;
      }
    } else {
      ar[0] += wi.re * wi.re + wi.im * wi.im;
    }
  }
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
  const int ncol = mode == 0 ? nv : 1;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    if (k < ncol) {
      const double sr = wsum(ar[k]);
      const double si = wsum(ai[k]);
      if (lane == 0) {
        red[wave][2 * k] = sr;
        red[wave][2 * k + 1] = si;
      }
    }
  }
  __syncthreads();
  double* p = part + ((long)leg * nbl + blockIdx.x) * (2 * NV);
  for (int k = threadIdx.x; k < 2 * ncol; k += 256)
    p[k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
}

// out[leg][k] = sum over blocks of part[leg][b][k], k < ncol2 (doubles). mode 0: coefficients;
// mode 1: sqrt of column 0 -> ab[leg][m + j] (beta); mode 2: coefficients and
// ab[leg][j] = out[leg][2j] (alpha, the real part of the j-th projection).
__global__ void __launch_bounds__(64)
    reduce_c_kernel(int nbl, int ncol2, int pstride, const double* __restrict__ part,
                    double* __restrict__ out, long sout, int mode, int j, int m,
                    double* __restrict__ ab, long sab) {
  const int leg = blockIdx.y, k = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nbl; b += 64) s += part[((long)leg * nbl + b) * pstride + k];
  s = wsum(s);
  if (threadIdx.x == 0) {
    if (mode == 1) {
      ab[leg * sab + m + j] = sqrt(s);
    } else {
      out[leg * sout + k] = s;
      if (mode == 2 && k == 2 * j) ab[leg * sab + j] = s;
    }
  }
}

// V[j+1] = w / max(beta, 1e-200) for every leg (complex vectors as 2n doubles)
__global__ void next_vec_kernel(long n2, const double* __restrict__ w, long sw2,
                                double* __restrict__ Vn, long sV2, const double* __restrict__ ab,
                                long sab, int m, int j) {
  const int leg = blockIdx.y;
  double d = ab[leg * sab + m + j];
  if (d < 1e-200) d = 1e-200;
  const double s = 1.0 / d;
  w += leg * sw2;
  Vn += leg * sV2;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n2; i += (long)gridDim.x * 256) Vn[i] = w[i] * s;
}

// out[leg] = sum_{k<nrows} c[leg][k] V[leg][k], complex coefficients on the device; with live
// extents (Ragged), a leg's entries outside them are zero in V and written as zeros unread
__global__ void lincomb_c_kernel(long n, int nrows, const c128* __restrict__ V, long ldv, long sV,
                                 const double* __restrict__ c, long sc, c128* __restrict__ out,
                                 long sout, Ragged rg) {
  const int leg = blockIdx.y;
  V += leg * sV;
  out += leg * sout;
  c += leg * sc;
  const int lx = rg.live ? rg.live[2 * leg] : 0, ly = rg.live ? rg.live[2 * leg + 1] : 0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    if (!live_entry(rg, lx, ly, i)) {
      out[i] = c128(0.0, 0.0);
      continue;
    }
    double re = 0.0, im = 0.0;
    for (int k = 0; k < nrows; ++k) {
      const c128 v = V[(long)k * ldv + i];
      const double cr = c[2 * k], ci = c[2 * k + 1];
      re += cr * v.re - ci * v.im;
      im += cr * v.im + ci * v.re;
    }
    out[i] = c128(re, im);
  }
}

template <int NV>
int cgs_pass(int batch, long n, int nv, const c128* V, long ldv, long sV, c128* w, long sw,
             const double* coef, long scoef, const double* beta, long sbeta, int beta_row, int mode,
             double* part, int nbl, const Ragged& rg, hipStream_t st) {
  hipLaunchKernelGGL(cgs_c_kernel<NV>, dim3(nbl, batch), dim3(256), 0, st, n, nv, V, ldv, sV, w, sw,
                     coef, scoef, beta, sbeta, beta_row, mode, part, nbl, rg);
  TNH_CHECK_LAUNCH("cgs_c");
  return TNH_OK;
}

int cgs_any(int batch, long n, int nv, const c128* V, long ldv, long sV, c128* w, long sw,
            const double* coef, long scoef, const double* beta, long sbeta, int beta_row, int mode,
            double* part, int nbl, const Ragged& rg, hipStream_t st) {
#define TNH_CGS(NV) \
  cgs_pass<NV>(batch, n, nv, V, ldv, sV, w, sw, coef, scoef, beta, sbeta, beta_row, mode, part, nbl, \
               rg, st)
  if (nv <= 4) return TNH_CGS(4);
  if (nv <= 8) return TNH_CGS(8);
  if (nv <= 16) return TNH_CGS(16);
  return TNH_CGS(32);
#undef TNH_CGS
}

}  // namespace
}  // namespace tnh

using namespace tnh;

extern "C" {

size_t tnh_expm_ws_bytes(int batch) {
  return align_up((size_t)kPartRows * 2 * kKMax * sizeof(double), 256) +
         2 * align_up((size_t)(batch > 0 ? batch : 1) * 2 * kKMax * sizeof(double), 256) + 256;
}

}  // extern "C"

namespace tnh {
namespace {
int expm_step_c(int batch, long n, int m, int j, void* V, long ldv, long sV, void* w, long sw,
                double* ab, long sab, void* ws, size_t ws_bytes, const Ragged& rg,
                tnh_stream_t stream) {
  TNH_REQUIRE(batch >= 1 && batch <= 65535 && n > 0 && ldv >= n, TNH_ERR_ARG,
              "tnh_expm_step_c: bad sizes (batch %d, n %ld, ldv %ld)", batch, n, ldv);
  TNH_REQUIRE(m >= 1 && m <= kKMax && j >= 0 && j < m, TNH_ERR_ARG,
              "tnh_expm_step_c: need 1 <= m <= %d and 0 <= j < m (m %d, j %d)", kKMax, m, j);
  TNH_REQUIRE(sab >= 2 * m, TNH_ERR_ARG, "tnh_expm_step_c: alpha/beta stride %ld < 2m", sab);
  const int nbl_max = kPartRows / batch;
  TNH_REQUIRE(nbl_max >= 1, TNH_ERR_ARG, "tnh_expm_step_c: batch %d above %d", batch, kPartRows);
  long nbl = (n + 255) / 256;
  if (nbl > nbl_max) nbl = nbl_max;
  Arena ar(ws, ws_bytes);
  double* part = ar.take<double>((size_t)kPartRows * 2 * kKMax);
  double* c1 = ar.take<double>((size_t)batch * 2 * kKMax);
  double* c2 = ar.take<double>((size_t)batch * 2 * kKMax);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "tnh_expm_step_c: workspace %zu < %zu bytes", ws_bytes, ar.used);
  hipStream_t st = static_cast<hipStream_t>(stream);
  c128* Vc = static_cast<c128*>(V);
  c128* wc = static_cast<c128*>(w);
  const int nv = j + 1;
  const int pst = 2 * (nv <= 4 ? 4 : nv <= 8 ? 8 : nv <= 16 ? 16 : 32);
  const long sc = 2 * kKMax;
#define TNH_T(x)                   \
  do {                             \
    int rc_ = (x);                 \
    if (rc_ != TNH_OK) return rc_; \
  } while (0)
  // three-term subtraction fused into pass 1's projection
  TNH_T(cgs_any(batch, n, nv, Vc, ldv, sV, wc, sw, nullptr, 0, j > 0 ? ab + m + (j - 1) : nullptr,
                sab, j - 1, 0, part, (int)nbl, rg, st));
  hipLaunchKernelGGL(reduce_c_kernel, dim3(2 * nv, batch), dim3(64), 0, st, (int)nbl, 2 * nv, pst,
                     part, c1, sc, 2, j, m, ab, sab);
  TNH_CHECK_LAUNCH("reduce_c");
  TNH_T(cgs_any(batch, n, nv, Vc, ldv, sV, wc, sw, c1, sc, nullptr, 0, 0, 0, part, (int)nbl, rg, st));
  hipLaunchKernelGGL(reduce_c_kernel, dim3(2 * nv, batch), dim3(64), 0, st, (int)nbl, 2 * nv, pst,
                     part, c2, sc, 0, j, m, ab, sab);
  TNH_CHECK_LAUNCH("reduce_c");
  TNH_T(cgs_any(batch, n, nv, Vc, ldv, sV, wc, sw, c2, sc, nullptr, 0, 0, 1, part, (int)nbl, rg, st));
  hipLaunchKernelGGL(reduce_c_kernel, dim3(1, batch), dim3(64), 0, st, (int)nbl, 1, pst, part,
                     c2, sc, 1, j, m, ab, sab);
  TNH_CHECK_LAUNCH("reduce_c");
  if (j + 1 < m) {
    long g = (2 * n + 255) / 256;
    if (g > 4096) g = 4096;
    hipLaunchKernelGGL(next_vec_kernel, dim3((unsigned)g, batch), dim3(256), 0, st, 2 * n,
                       reinterpret_cast<const double*>(wc), 2 * sw,
                       reinterpret_cast<double*>(Vc + (long)(j + 1) * ldv), 2 * sV, ab, sab, m, j);
    TNH_CHECK_LAUNCH("next_vec");
  }
#undef TNH_T
  return TNH_OK;
}
}  // namespace
}  // namespace tnh

extern "C" {

int tnh_expm_step_c(int batch, long n, int m, int j, void* V, long ldv, long sV, void* w, long sw,
                    double* ab, long sab, void* ws, size_t ws_bytes, tnh_stream_t stream) {
  return tnh::expm_step_c(batch, n, m, j, V, ldv, sV, w, sw, ab, sab, ws, ws_bytes,
                          tnh::Ragged{nullptr, 1, 1}, stream);
}

int tnh_expm_step_c_ragged(int batch, long n, int m, int j, void* V, long ldv, long sV, void* w,
                           long sw, double* ab, long sab, const int* live, int rowlen, int bpad,
                           void* ws, size_t ws_bytes, tnh_stream_t stream) {
  TNH_REQUIRE(live && rowlen >= 1 && bpad >= 1 && rowlen % bpad == 0 && n % rowlen == 0 &&
                  n < (1L << 31),
              TNH_ERR_ARG, "tnh_expm_step_c_ragged: bad layout (n %ld, rowlen %d, bpad %d)", n,
              rowlen, bpad);
  return tnh::expm_step_c(batch, n, m, j, V, ldv, sV, w, sw, ab, sab, ws, ws_bytes,
                          tnh::Ragged{live, rowlen, bpad}, stream);
}

int tnh_lincomb_c(int batch, long n, int nrows, const void* V, long ldv, long sV, const double* c,
                  long sc, void* out, long sout, tnh_stream_t stream) {
  return tnh_lincomb_c_ragged(batch, n, nrows, V, ldv, sV, c, sc, out, sout, nullptr, 1, 1,
                              stream);
}

int tnh_lincomb_c_ragged(int batch, long n, int nrows, const void* V, long ldv, long sV,
                         const double* c, long sc, void* out, long sout, const int* live,
                         int rowlen, int bpad, tnh_stream_t stream) {
  TNH_REQUIRE(batch >= 1 && batch <= 65535 && n >= 0 && nrows >= 0 && (nrows == 0 || ldv >= n),
              TNH_ERR_ARG, "tnh_lincomb_c: bad sizes");
  TNH_REQUIRE(!live || (rowlen >= 1 && bpad >= 1 && rowlen % bpad == 0 && n % rowlen == 0 &&
                        n < (1L << 31)),
              TNH_ERR_ARG, "tnh_lincomb_c_ragged: bad layout (n %ld, rowlen %d, bpad %d)", n,
              rowlen, bpad);
  if (n == 0) return TNH_OK;
  long g = (n + 255) / 256;
  if (g > 4096) g = 4096;
  hipLaunchKernelGGL(lincomb_c_kernel, dim3((unsigned)g, batch), dim3(256), 0,
                     static_cast<hipStream_t>(stream), n, nrows, static_cast<const c128*>(V), ldv, sV,
                     c, sc, static_cast<c128*>(out), sout, Ragged{live, rowlen, bpad});
  TNH_CHECK_LAUNCH("lincomb_c");
  return TNH_OK;
}

}  // extern "C"
