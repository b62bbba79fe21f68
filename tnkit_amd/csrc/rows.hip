// Row projections and row combinations of a stack of vectors, in any of the path's element types:
//   tnh_rows_dot:     out[r] = sum_i V[r][i] x[i]          (einsum "kn,n->k", no conjugation)
//   tnh_rows_combine: out[i] = sum_r c[r] V[r][i]          (einsum "k,kn->n", c on the device)
// These are the Krylov projections of tnkit's eigensolvers as the backend sees them
// (tnkit/eigensolver.py:160-161, :567, :588-595, :672). Both stream V once; the dot is a
// deterministic two-stage reduction (one partial per workgroup and row, summed in block order).
#include "cplx.h"
#include "tnh_common.h"

namespace tnh {
namespace {

constexpr int kRowsBlocks = 1024;  // most workgroups of a dot: rows of the partial buffer
constexpr int kRowsChunk = 32;     // rows per dot launch (accumulators per thread)

template <typename T>
struct acc_of {
  typedef double type;
};
template <>
struct acc_of<c128> {
  typedef c128 type;
};
template <>
struct acc_of<c64> {
  typedef c128 type;
};

template <typename A, typename T>
__device__ __forceinline__ A widen(T v) {
  return static_cast<A>(v);
}
template <>
__device__ __forceinline__ c128 widen<c128, c64>(c64 v) {
  return c128(v);
}

template <typename T>
__device__ __forceinline__ T narrow_to(double v) {
  return static_cast<T>(v);
}
template <typename T>
__device__ __forceinline__ T narrow_to(c128 v) {
  return T(v);
}

__device__ __forceinline__ double wsum(double v) {
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
  return v;
}
__device__ __forceinline__ c128 wsum(c128 v) { return c128(wsum(v.re), wsum(v.im)); }

// part[blockIdx.x * kRowsChunk + r] = this block's sum of V[r][i] x[i], r < nr
template <typename T>
__global__ void __launch_bounds__(256)
    rows_dot_kernel(long n, int nr, const T* __restrict__ V, long ldv, const T* __restrict__ x,
                    typename acc_of<T>::type* __restrict__ part) {
  typedef typename acc_of<T>::type A;
  A acc[kRowsChunk];
#pragma unroll
  for (int r = 0; r < kRowsChunk; ++r) acc[r] = A(0.0);
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    const A xi = widen<A>(x[i]);
#pragma unroll
    for (int r = 0; r < kRowsChunk; ++r)
      if (r < nr) acc[r] += widen<A>(V[(long)r * ldv + i]) * xi;
  }
  // raw storage: the complex type has a constructor, which __shared__ arrays may not
  __shared__ __align__(16) unsigned char red_buf[4 * kRowsChunk * sizeof(A)];
  A* red = reinterpret_cast<A*>(red_buf);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
#pragma unroll
  for (int r = 0; r < kRowsChunk; ++r) {
    if (r < nr) {
      const A s = wsum(acc[r]);
      if (lane == 0) red[wave * kRowsChunk + r] = s;
    }
  }
  __syncthreads();
  for (int r = threadIdx.x; r < nr; r += blockDim.x)
    part[(long)blockIdx.x * kRowsChunk + r] = red[r] + red[kRowsChunk + r] +
                                              red[2 * kRowsChunk + r] + red[3 * kRowsChunk + r];
}

// out[r] = sum_b part[b][r] in block order; one wave per row
template <typename T>
__global__ void __launch_bounds__(64)
    rows_reduce_kernel(int nb, const typename acc_of<T>::type* __restrict__ part,
                       T* __restrict__ out) {
  typedef typename acc_of<T>::type A;
  const int r = blockIdx.x;
  A s = A(0.0);
  for (int b = threadIdx.x; b < nb; b += 64) s += part[(long)b * kRowsChunk + r];
  s = wsum(s);
  if (threadIdx.x == 0) out[r] = narrow_to<T>(s);
}

template <typename T>
__global__ void __launch_bounds__(256)
    rows_combine_kernel(long n, int k, const T* __restrict__ c, const T* __restrict__ V,
                        long ldv, T* __restrict__ out) {
  typedef typename acc_of<T>::type A;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < n; i += (long)gridDim.x * 256) {
    A s = A(0.0);
    for (int r = 0; r < k; ++r) s += widen<A>(c[r]) * widen<A>(V[(long)r * ldv + i]);
    out[i] = narrow_to<T>(s);
  }
}

inline int blocks_for(long n, int cap) {
  long b = (n + 255) / 256;
  if (b < 1) b = 1;
  if (b > cap) b = cap;
  return static_cast<int>(b);
}

template <typename T>
int rows_dot(int k, long n, const void* V, long ldv, const void* x, void* out, void* ws,
             size_t ws_bytes, hipStream_t st) {
  typedef typename acc_of<T>::type A;
  Arena ar(ws, ws_bytes);
  A* part = ar.take<A>((size_t)kRowsBlocks * kRowsChunk);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "tnh_rows_dot: workspace %zu < %zu bytes", ws_bytes, ar.used);
  const int nb = blocks_for(n, kRowsBlocks);
  for (int r0 = 0; r0 < k; r0 += kRowsChunk) {
    const int nr = k - r0 < kRowsChunk ? k - r0 : kRowsChunk;
    hipLaunchKernelGGL(rows_dot_kernel<T>, dim3(nb), dim3(256), 0, st, n, nr,
                       static_cast<const T*>(V) + (long)r0 * ldv, ldv, static_cast<const T*>(x),
                       part);
    TNH_CHECK_LAUNCH("rows_dot");
    hipLaunchKernelGGL(rows_reduce_kernel<T>, dim3(nr), dim3(64), 0, st, nb, part,
                       static_cast<T*>(out) + r0);
    TNH_CHECK_LAUNCH("rows_reduce");
  }
  return TNH_OK;
}

template <typename T>
int rows_combine(int k, long n, const void* c, const void* V, long ldv, void* out,
                 hipStream_t st) {
  hipLaunchKernelGGL(rows_combine_kernel<T>, dim3(blocks_for(n, 16384)), dim3(256), 0, st, n, k,
                     static_cast<const T*>(c), static_cast<const T*>(V), ldv,
                     static_cast<T*>(out));
  TNH_CHECK_LAUNCH("rows_combine");
  return TNH_OK;
}

}  // namespace
}  // namespace tnh

using namespace tnh;

extern "C" {

size_t tnh_rows_ws_bytes(int dtype) {
  (void)dtype;
  return (size_t)kRowsBlocks * kRowsChunk * sizeof(c128) + 256;
}

int tnh_rows_dot(int dtype, int k, long n, const void* V, long ldv, const void* x, void* out,
                 void* ws, size_t ws_bytes, tnh_stream_t stream) {
  TNH_REQUIRE(k >= 0 && n >= 0 && (k <= 1 || ldv >= n), TNH_ERR_ARG,
              "tnh_rows_dot: bad sizes k=%d n=%ld ldv=%ld", k, n, ldv);
  if (k == 0) return TNH_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case TNH_F64: return rows_dot<double>(k, n, V, ldv, x, out, ws, ws_bytes, st);
    case TNH_F32: return rows_dot<float>(k, n, V, ldv, x, out, ws, ws_bytes, st);
    case TNH_C128: return rows_dot<c128>(k, n, V, ldv, x, out, ws, ws_bytes, st);
    case TNH_C64: return rows_dot<c64>(k, n, V, ldv, x, out, ws, ws_bytes, st);
  }
  set_error("tnh_rows_dot: unsupported dtype %d", dtype);
  return TNH_ERR_UNSUPPORTED;
}

int tnh_rows_combine(int dtype, int k, long n, const void* c, const void* V, long ldv, void* out,
                     tnh_stream_t stream) {
  TNH_REQUIRE(k >= 0 && n >= 0 && (k <= 1 || ldv >= n), TNH_ERR_ARG,
              "tnh_rows_combine: bad sizes k=%d n=%ld ldv=%ld", k, n, ldv);
  if (n == 0) return TNH_OK;
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case TNH_F64: return rows_combine<double>(k, n, c, V, ldv, out, st);
    case TNH_F32: return rows_combine<float>(k, n, c, V, ldv, out, st);
    case TNH_C128: return rows_combine<c128>(k, n, c, V, ldv, out, st);
    case TNH_C64: return rows_combine<c64>(k, n, c, V, ldv, out, st);
  }
  set_error("tnh_rows_combine: unsupported dtype %d", dtype);
  return TNH_ERR_UNSUPPORTED;
}

}  // extern "C"
