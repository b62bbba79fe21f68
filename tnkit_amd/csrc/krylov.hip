// Float64 Krylov-vector kernels: the Lanczos three-term / arrowhead subtraction, alpha, the two
// classical Gram-Schmidt passes and beta of one thick-restart Lanczos iteration
// (tnkit/eigensolver.py:560-604), plus linear combinations of basis rows (the Ritz vector,
// tnkit/eigensolver.py:670-673) and the small helpers around them.
//
// Every reduction is two-stage and deterministic: each workgroup writes one partial per
// coefficient, and a second kernel sums the partials in a fixed order. The basis V is read once
// per Gram-Schmidt pass: a thread loads V[0..nv)[i] into registers, subtracts the previous
// pass's projection from w[i] and accumulates the next pass's dot products from the same
// registers, so one iteration reads V three times (the reference form reads it four times).
#include "tnh_common.h"

namespace tnh {

constexpr int kHostCoef = 128;  // coefficients passed by value (Krylov windows up to 127 rows)

struct HostCoef {
  double c[kHostCoef];
};

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// Sums `vals[0..NV)` over the block and writes them to part[blockIdx.x*ld + k].
template <int NV>
__device__ __forceinline__ void block_partials(double (&vals)[NV], int nv, double* part, int ld) {
  __shared__ double red[4][NV];
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int k = 0; k < NV; ++k) {
    if (k < nv) {
      const double s = wave_sum(vals[k]);
      if (lane == 0) red[wave][k] = s;
    }
  }
  __syncthreads();
  for (int k = threadIdx.x; k < nv; k += blockDim.x)
    part[(long)blockIdx.x * ld + k] = red[0][k] + red[1][k] + red[2][k] + red[3][k];
}

// w[i] -= sum_r hc[r] * V[sub_base + r][i];   partial <V[j], w>
__global__ void __launch_bounds__(256)
    orth_pre_kernel(long n, const double* __restrict__ V, long ldv, double* __restrict__ w,
                    int nsub, int sub_base, HostCoef hc, int j, double* __restrict__ part) {
  double acc[1] = {0.0};
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    double wi = w[i];
    for (int r = 0; r < nsub; ++r) wi -= hc.c[r] * V[(long)(sub_base + r) * ldv + i];
    if (nsub > 0) w[i] = wi;
    acc[0] += V[(long)j * ldv + i] * wi;
  }
  block_partials<1>(acc, 1, part, 1);
}

// w[i] -= sum_{k<nv} dc[k] * V[k][i]; then partial dots <V[k], w> (mode 0), or partial |w|^2
// (mode 1), or nothing (mode 2). V rows are held in registers between the subtraction and the
// dots. The rows are read through one buffer resource: the element offset is one VGPR shared by
// every row and the row offset k * ldv * 8 is a scalar, so no per-row 64-bit address takes
// registers (the caller guarantees every offset fits in 32 bits); the coefficients are uniform
// loads into scalar registers. At 41 rows this holds two waves per SIMD and reads the basis at
// 5.3 TB/s (n = 2^22), against 4.6 TB/s for per-lane 64-bit addresses at one wave per SIMD.
template <int NV>
__global__ void __launch_bounds__(256, NV <= 56 ? 2 : 1)
    orth_pass_kernel(unsigned n, int nv, const double* __restrict__ V, long ldv,
                     double* __restrict__ w, const double* __restrict__ dc, int mode,
                     double* __restrict__ part) {
  const __amdgpu_buffer_rsrc_t rs =
      __builtin_amdgcn_make_buffer_rsrc(const_cast<double*>(V), (short)0, -1, 0x00020000);
  const unsigned rowb = (unsigned)(ldv * 8);
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (unsigned i = blockIdx.x * 256u + threadIdx.x; i < n; i += gridDim.x * 256u) {
    double v[NV];
    double wi = w[i];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      // no branch per row: rows past nv load the last live row and weigh zero, so every load
      // of the element is in flight before the first subtraction waits
      const int kk = k < nv ? k : nv - 1;
      const double x = __builtin_bit_cast(
          double, __builtin_amdgcn_raw_buffer_load_b64(rs, i * 8u, (unsigned)kk * rowb, 0));
      const double ck = dc[kk];
      v[k] = k < nv ? x : 0.0;
      wi -= (k < nv ? ck : 0.0) * x;
    }
    w[i] = wi;
    if (mode == 0) {
#pragma unroll
      for (int k = 0; k < NV; ++k) acc[k] += v[k] * wi;
    } else if (mode == 1) {
      acc[0] += wi * wi;
    }
  }
  if (mode == 2) return;  // subtraction only (uniform over the block, before any barrier)
  block_partials<NV>(acc, mode == 0 ? nv : 1, part, mode == 0 ? nv : 1);
}

// The same pass for bases whose byte offsets do not fit in 32 bits: per-lane addresses.
template <int NV>
__global__ void __launch_bounds__(256)
    orth_pass_wide_kernel(long n, int nv, const double* __restrict__ V, long ldv,
                          double* __restrict__ w, const double* __restrict__ dc, int mode,
                          double* __restrict__ part) {
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    double v[NV];
    double wi = w[i];
#pragma unroll
    for (int k = 0; k < NV; ++k) {
      const int kk = k < nv ? k : nv - 1;
      const double x = V[(long)kk * ldv + i];
      const double ck = dc[kk];
      v[k] = k < nv ? x : 0.0;
      wi -= (k < nv ? ck : 0.0) * x;
    }
    w[i] = wi;
    if (mode == 0) {
#pragma unroll
      for (int k = 0; k < NV; ++k) acc[k] += v[k] * wi;
    } else if (mode == 1) {
      acc[0] += wi * wi;
    }
  }
  if (mode == 2) return;
  block_partials<NV>(acc, mode == 0 ? nv : 1, part, mode == 0 ? nv : 1);
}

// partial dots <V[k], x> for k < nv (x is only read)
template <int NV>
__global__ void __launch_bounds__(256)
    dots_kernel(long n, int nv, const double* __restrict__ V, long ldv, const double* __restrict__ x,
                double* __restrict__ part) {
  double acc[NV];
#pragma unroll
  for (int k = 0; k < NV; ++k) acc[k] = 0.0;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const double xi = x[i];
#pragma unroll
    for (int k = 0; k < NV; ++k)
      if (k < nv) acc[k] += V[(long)k * ldv + i] * xi;
  }
  block_partials<NV>(acc, nv, part, nv);
}

// out[k] = sum_b part[b*ld + k]  for k < ncol, in block order. mode 0: plain; 1: sqrt (norms);
// 2: alpha -- scal[0] = sum and coef[k] = (k == jsel) ? sum : 0 for k < ncoef; 3: scale * sum.
__global__ void __launch_bounds__(256)
    reduce_kernel(int nb, int ncol, const double* __restrict__ part, int ld, double* __restrict__ out,
                  int mode, int jsel, int ncoef, double* __restrict__ coef, double scale) {
  __shared__ double red[4];
  const int k = blockIdx.x;
  double s = 0.0;
  for (int b = threadIdx.x; b < nb; b += blockDim.x) s += part[(long)b * ld + k];
  s = wave_sum(s);
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = s;
  __syncthreads();
  if (threadIdx.x == 0) {
    double tot = red[0] + red[1] + red[2] + red[3];
    if (mode == 1) tot = sqrt(tot);
    if (mode == 3) tot *= scale;
    out[k] = tot;
    if (mode == 2)
      for (int r = 0; r < ncoef; ++r) coef[r] = (r == jsel) ? tot : 0.0;
  }
}

// out[i] = beta * out[i] + sum_r hc[r] * V[r][i]
__global__ void __launch_bounds__(256)
    lincomb_kernel(long n, int nrows, const double* __restrict__ V, long ldv, HostCoef hc,
                   double beta, double* __restrict__ out) {
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    double s = (beta != 0.0) ? beta * out[i] : 0.0;
    for (int r = 0; r < nrows; ++r) s += hc.c[r] * V[(long)r * ldv + i];
    out[i] = s;
  }
}

// y[i] = x[i] * (mul / div[0])  -- division by a device scalar (beta), with tnkit's
// guard_denom clamp: a divisor below 1e-200 is replaced by 1e-200 (tnkit/backend.py:699-708)
// Elementwise y[i] = f(x[i]): each thread moves 4 consecutive elements per step (two 16-byte
// loads in flight for doubles) so that enough bytes are in flight to cover HBM latency.
template <typename TO, typename TI, typename F>
__device__ __forceinline__ void map4(long n, const TI* __restrict__ x, TO* __restrict__ y, F f) {
  const long stride = (long)gridDim.x * 256 * 4;
  for (long i0 = (blockIdx.x * 256L + threadIdx.x) * 4; i0 < n; i0 += stride) {
    if (i0 + 3 < n) {
      TI a0 = x[i0], a1 = x[i0 + 1], a2 = x[i0 + 2], a3 = x[i0 + 3];
      y[i0] = f(a0);
      y[i0 + 1] = f(a1);
      y[i0 + 2] = f(a2);
      y[i0 + 3] = f(a3);
    } else {
      for (long i = i0; i < n; ++i) y[i] = f(x[i]);
    }
  }
}

__global__ void scale_by_kernel(long n, const double* __restrict__ x, const double* __restrict__ div,
                                double mul, double* __restrict__ y) {
  double d = div ? div[0] : 1.0;
  if (div && d < 1e-200) d = 1e-200;
  const double s = mul / d;
  map4(n, x, y, [s](double v) { return v * s; });
}

template <typename TO, typename TI>
__global__ void convert_kernel(long n, const TI* __restrict__ x, TO* __restrict__ y) {
  map4(n, x, y, [](TI v) { return static_cast<TO>(v); });
}

namespace {

// workgroups of the reducing kernels: each writes one partial per coefficient, so their number
// is capped at the partial buffer's 1024 rows (four per CU)
inline int nblocks(long n) {
  long b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 1024) b = 1024;
  return static_cast<int>(b);
}

inline int nblocks_lin(long n) {
  long b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > 16384) b = 16384;
  return static_cast<int>(b);
}

// workgroups of the elementwise kernels (4 elements per thread per step)
inline int nblocks_stream(long n) {
  long b = (n + 1023) / 1024;
  if (b < 1) b = 1;
  if (b > 16384) b = 16384;
  return static_cast<int>(b);
}

// The buffer-resource pass addresses V[k][i] as k * ldv * 8 + i * 8 in 32-bit offsets.
inline bool offsets_fit(long n, int nv, long ldv) {
  const unsigned long long span =
      (unsigned long long)(nv > 0 ? nv - 1 : 0) * (unsigned long long)ldv * 8ull +
      (unsigned long long)n * 8ull;
  return span < (1ull << 32) - 64;
}

// The pass kernels read rows 0..nv-1 and coefficients dc[0..nv-1] with no per-row branch, so
// nv >= 1 and dc != nullptr are required here.
template <int NV>
int pass(long n, int nv, const double* V, long ldv, double* w, const double* dc, int mode,
         double* part, int nb, hipStream_t st) {
  TNH_REQUIRE(nv >= 1 && nv <= NV && dc != nullptr && V != nullptr, TNH_ERR_ARG,
              "krylov: a pass needs 1..%d basis rows and their coefficients (nv=%d)", NV, nv);
  if (offsets_fit(n, nv, ldv))
    hipLaunchKernelGGL(orth_pass_kernel<NV>, dim3(nb), dim3(256), 0, st, (unsigned)n, nv, V, ldv,
                       w, dc, mode, part);
  else
    hipLaunchKernelGGL(orth_pass_wide_kernel<NV>, dim3(nb), dim3(256), 0, st, n, nv, V, ldv, w,
                       dc, mode, part);
  TNH_CHECK_LAUNCH("orth_pass");
  return TNH_OK;
}

// Row counts come in steps of four, so a pass carries at most three idle rows of registers.
int pass_any(long n, int nv, const double* V, long ldv, double* w, const double* dc, int mode,
             double* part, int nb, hipStream_t st) {
  TNH_REQUIRE(nv >= 1, TNH_ERR_ARG, "krylov: a pass needs at least one basis row, got %d", nv);
  switch ((nv + 3) / 4) {
    case 1: return pass<4>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 2: return pass<8>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 3: return pass<12>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 4: return pass<16>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 5: return pass<20>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 6: return pass<24>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 7: return pass<28>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 8: return pass<32>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 9: return pass<36>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 10: return pass<40>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 11: return pass<44>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 12: return pass<48>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 13:
    case 14: return pass<56>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    case 15:
    case 16: return pass<64>(n, nv, V, ldv, w, dc, mode, part, nb, st);
    default: break;
  }
  set_error("krylov: at most 64 live basis vectors are supported, got %d", nv);
  return TNH_ERR_ARG;
}

template <int NV>
int dots(long n, int nv, const double* V, long ldv, const double* x, double* part, int nb,
         hipStream_t st) {
  hipLaunchKernelGGL(dots_kernel<NV>, dim3(nb), dim3(256), 0, st, n, nv, V, ldv, x, part);
  TNH_CHECK_LAUNCH("dots");
  return TNH_OK;
}

int dots_any(long n, int nv, const double* V, long ldv, const double* x, double* part, int nb,
             hipStream_t st) {
  if (nv <= 4) return dots<4>(n, nv, V, ldv, x, part, nb, st);
  if (nv <= 8) return dots<8>(n, nv, V, ldv, x, part, nb, st);
  if (nv <= 16) return dots<16>(n, nv, V, ldv, x, part, nb, st);
  if (nv <= 32) return dots<32>(n, nv, V, ldv, x, part, nb, st);
  if (nv <= 64) return dots<64>(n, nv, V, ldv, x, part, nb, st);
  set_error("rank_update: at most 64 vectors are supported, got %d", nv);
  return TNH_ERR_ARG;
}

int reduce(int nb, int ncol, const double* part, int ld, double* out, int mode, int jsel, int ncoef,
           double* coef, hipStream_t st, double scale = 1.0) {
  hipLaunchKernelGGL(reduce_kernel, dim3(ncol), dim3(256), 0, st, nb, ncol, part, ld, out, mode, jsel,
                     ncoef, coef, scale);
  TNH_CHECK_LAUNCH("reduce");
  return TNH_OK;
}

#define TNH_TRY(x)                 \
  do {                             \
    int rc_ = (x);                 \
    if (rc_ != TNH_OK) return rc_; \
  } while (0)

}  // namespace
}  // namespace tnh

using namespace tnh;

extern "C" {

size_t tnh_krylov_ws_bytes(long n, int nmax) {
  (void)n;
  // partials: 1024 blocks x 64 columns, plus three coefficient vectors
  return align_up(1024 * 64 * sizeof(double), 256) + 3 * align_up(64 * sizeof(double), 256) +
         256 + (size_t)0 * nmax;
}

int tnh_krylov_orth(long n, int nlive, const double* V, long ldv, double* w, int j, int nsub,
                    int sub_base, const double* sub_coef, double* scal, void* ws, size_t ws_bytes,
                    tnh_stream_t stream) {
  TNH_REQUIRE(n > 0 && ldv >= n, TNH_ERR_ARG, "tnh_krylov_orth: bad n=%ld ldv=%ld", n, ldv);
  TNH_REQUIRE(nlive >= 1 && nlive <= 64 && j >= 0 && j < nlive, TNH_ERR_ARG,
              "tnh_krylov_orth: need 1 <= nlive <= 64 and 0 <= j < nlive (nlive=%d, j=%d)", nlive, j);
  TNH_REQUIRE(nsub >= 0 && nsub < kHostCoef && sub_base >= 0, TNH_ERR_ARG,
              "tnh_krylov_orth: nsub=%d out of range", nsub);
  HostCoef hc;
  memset(&hc, 0, sizeof(hc));
  for (int r = 0; r < nsub; ++r) hc.c[r] = sub_coef[r];
  Arena ar(ws, ws_bytes);
  double* part = ar.take<double>(1024 * 64);
  double* c0 = ar.take<double>(64);
  double* c1 = ar.take<double>(64);
  double* c2 = ar.take<double>(64);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "tnh_krylov_orth: workspace %zu < %zu bytes", ws_bytes, ar.used);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nb = nblocks(n);
  // w -= sub; alpha = <V[j], w>
  hipLaunchKernelGGL(orth_pre_kernel, dim3(nb), dim3(256), 0, st, n, V, ldv, w, nsub, sub_base, hc,
                     j, part);
  TNH_CHECK_LAUNCH("orth_pre");
  TNH_TRY(reduce(nb, 1, part, 1, scal + 0, 2, j, nlive, c0, st));
  // w -= alpha V[j]; c1 = V w
  TNH_TRY(pass_any(n, nlive, V, ldv, w, c0, 0, part, nb, st));
  TNH_TRY(reduce(nb, nlive, part, nlive, c1, 0, 0, 0, nullptr, st));
  // w -= c1 V; c2 = V w
  TNH_TRY(pass_any(n, nlive, V, ldv, w, c1, 0, part, nb, st));
  TNH_TRY(reduce(nb, nlive, part, nlive, c2, 0, 0, 0, nullptr, st));
  // w -= c2 V; beta = |w|
  TNH_TRY(pass_any(n, nlive, V, ldv, w, c2, 1, part, nb, st));
  TNH_TRY(reduce(nb, 1, part, 1, scal + 1, 1, 0, 0, nullptr, st));
  return TNH_OK;
}

int tnh_rank_update(long n, int k, const double* V, long ldv, const double* x, double weight,
                    double* y, void* ws, size_t ws_bytes, tnh_stream_t stream) {
  TNH_REQUIRE(n > 0 && ldv >= n, TNH_ERR_ARG, "tnh_rank_update: bad n=%ld ldv=%ld", n, ldv);
  TNH_REQUIRE(k >= 0 && k <= 64, TNH_ERR_ARG, "tnh_rank_update: need 0 <= k <= 64, got %d", k);
  if (k == 0) return TNH_OK;
  Arena ar(ws, ws_bytes);
  double* part = ar.take<double>(1024 * 64);
  double* c = ar.take<double>(64);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "tnh_rank_update: workspace %zu < %zu bytes", ws_bytes, ar.used);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nb = nblocks(n);
  // c = -weight * V x  (every coefficient is complete before y is touched, so x may alias y)
  TNH_TRY(dots_any(n, k, V, ldv, x, part, nb, st));
  TNH_TRY(reduce(nb, k, part, k, c, 3, 0, 0, nullptr, st, -weight));
  // y -= c V
  return pass_any(n, k, V, ldv, y, c, 2, part, nb, st);
}

int tnh_lincomb(long n, int nrows, const double* V, long ldv, const double* coef, double beta,
                double* out, tnh_stream_t stream) {
  TNH_REQUIRE(n >= 0 && nrows >= 0 && nrows < kHostCoef && (nrows == 0 || ldv >= n), TNH_ERR_ARG,
              "tnh_lincomb: bad sizes n=%ld nrows=%d ldv=%ld", n, nrows, ldv);
  if (n == 0) return TNH_OK;
  HostCoef hc;
  memset(&hc, 0, sizeof(hc));
  for (int r = 0; r < nrows; ++r) hc.c[r] = coef[r];
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(lincomb_kernel, dim3(nblocks_lin(n)), dim3(256), 0, st, n, nrows, V, ldv, hc, beta,
                     out);
  TNH_CHECK_LAUNCH("lincomb");
  return TNH_OK;
}

int tnh_norm(long n, const double* x, double* out, void* ws, size_t ws_bytes, tnh_stream_t stream) {
  TNH_REQUIRE(n > 0, TNH_ERR_ARG, "tnh_norm: n must be positive");
  Arena ar(ws, ws_bytes);
  double* part = ar.take<double>(1024 * 64);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "tnh_norm: workspace %zu < %zu bytes", ws_bytes, ar.used);
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nb = nblocks(n);
  HostCoef hc;
  memset(&hc, 0, sizeof(hc));
  // the dot of x with itself (no subtraction, so x is only read), then the square root
  hipLaunchKernelGGL(orth_pre_kernel, dim3(nb), dim3(256), 0, st, n, x, n, const_cast<double*>(x),
                     0, 0, hc, 0, part);
  TNH_CHECK_LAUNCH("norm");
  return reduce(nb, 1, part, 1, out, 1, 0, 0, nullptr, st);
}

int tnh_dot(long n, const double* x, const double* y, double* out, void* ws, size_t ws_bytes,
            tnh_stream_t stream) {
  TNH_REQUIRE(n > 0, TNH_ERR_ARG, "tnh_dot: n must be positive");
  Arena ar(ws, ws_bytes);
  double* part = ar.take<double>(1024 * 64);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "tnh_dot: workspace %zu < %zu bytes", ws_bytes, ar.used);
  HostCoef hc;
  memset(&hc, 0, sizeof(hc));
  hipStream_t st = static_cast<hipStream_t>(stream);
  const int nb = nblocks(n);
  hipLaunchKernelGGL(orth_pre_kernel, dim3(nb), dim3(256), 0, st, n, x, n, const_cast<double*>(y), 0,
                     0, hc, 0, part);
  TNH_CHECK_LAUNCH("dot");
  return reduce(nb, 1, part, 1, out, 0, 0, 0, nullptr, st);
}

int tnh_scale_by(long n, const double* x, const double* div, double mul, double* y,
                 tnh_stream_t stream) {
  if (n <= 0) return TNH_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipLaunchKernelGGL(scale_by_kernel, dim3(nblocks_stream(n)), dim3(256), 0, st, n, x, div, mul, y);
  TNH_CHECK_LAUNCH("scale_by");
  return TNH_OK;
}

int tnh_convert(int dst_dtype, int src_dtype, long n, const void* src, void* dst,
                tnh_stream_t stream) {
  if (n <= 0) return TNH_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dst_dtype == TNH_F32 && src_dtype == TNH_F64) {
    hipLaunchKernelGGL((convert_kernel<float, double>), dim3(nblocks_stream(n)), dim3(256), 0, st, n,
                       static_cast<const double*>(src), static_cast<float*>(dst));
  } else if (dst_dtype == TNH_F64 && src_dtype == TNH_F32) {
    hipLaunchKernelGGL((convert_kernel<double, float>), dim3(nblocks_stream(n)), dim3(256), 0, st, n,
                       static_cast<const float*>(src), static_cast<double*>(dst));
  } else if (dst_dtype == TNH_BF16 && src_dtype == TNH_F64) {  // one rounding, to nearest even
    hipLaunchKernelGGL((convert_kernel<__bf16, double>), dim3(nblocks_stream(n)), dim3(256), 0, st,
                       n, static_cast<const double*>(src), static_cast<__bf16*>(dst));
  } else {
    set_error("tnh_convert: unsupported pair (%d <- %d)", dst_dtype, src_dtype);
    return TNH_ERR_UNSUPPORTED;
  }
  TNH_CHECK_LAUNCH("convert");
  return TNH_OK;
}

}  // extern "C"
