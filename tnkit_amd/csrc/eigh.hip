// Batched Hermitian eigensolver for small matrices (n <= 64): one workgroup per matrix, the
// matrix and its eigenvector accumulator held in LDS for the whole solve, parallel cyclic Jacobi.
//
// A sweep visits every index pair once in n2 - 1 rounds of n2 / 2 disjoint pairs (round-robin
// tournament order, n2 = n rounded up to even; an odd n gets one zero pad index whose rotations
// are the identity). In a round each pair (p, q) gets the complex Jacobi rotation
//   U = [[c, s], [-s e, c e]],  e = conj(A_pq) / |A_pq|,
// which makes the 2x2 block [[A_pp, A_pq], [A_qp, A_qq]] diagonal (a real symmetric rotation
// after the phase diag(1, e)). Because the pairs are disjoint, A <- U^H A U splits into
// independent 2x2 blocks (rows of pair k, columns of pair l), one per work item, and V <- V U
// into independent row pieces. Rotations with |A_pq| <= 1e-16 sqrt(|A_pp A_qq|) are skipped.
// A solve stops after the first sweep that applies no rotation or ends with an off-diagonal
// Frobenius norm <= tol * ||A||_F.
//
// Used for the Gram matrices of the truncation (tnkit/linalg.py:138-159, the batched form of
// tnkit/batch_tdvp.py:572-599), where rocSOLVER's batched eigh spends milliseconds per call on
// matrices this size.
#include "cplx.h"
#include "tnh_common.h"

namespace tnh {

constexpr int kJacMax = 64;
constexpr int kJacThreads = 256;

__device__ __forceinline__ double jre(double x) { return x; }
__device__ __forceinline__ double jre(c128 x) { return x.re; }
__device__ __forceinline__ double jabs2(double x) { return x * x; }
__device__ __forceinline__ double jabs2(c128 x) { return x.re * x.re + x.im * x.im; }
__device__ __forceinline__ double jconj(double x) { return x; }
__device__ __forceinline__ c128 jconj(c128 x) { return conj(x); }
__device__ __forceinline__ double jscale(double s, double x) { return s * x; }
__device__ __forceinline__ c128 jscale(double s, c128 x) { return {s * x.re, s * x.im}; }
__device__ __forceinline__ double jmul(double a, double b) { return a * b; }
__device__ __forceinline__ c128 jmul(c128 a, c128 b) { return a * b; }

template <typename T>
struct JacobiShared {
  T A[kJacMax * kJacMax];
  T V[kJacMax * kJacMax];
  T e[kJacMax / 2];
  double c[kJacMax / 2], s[kJacMax / 2], dp[kJacMax / 2], dq[kJacMax / 2];
  int p[kJacMax / 2], q[kJacMax / 2], on[kJacMax / 2];
  int rank[kJacMax];
  double red[kJacThreads / kWave];
  int nrot;
};

__device__ __forceinline__ double jac_block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < kJacThreads / kWave; ++w) t += red[w];
  return t;
}

// matrix b: A_b = A + b*sA (row i at i*lda), plus add_b[j] on the diagonal when add is set;
// w_b ascending eigenvalues, Z_b[i*ldz + j] = component i of eigenvector j (Z may alias A).
// sweeps[b] (optional) = sweeps run, negated when max_sweeps ended the solve unconverged.
template <typename T>
__global__ void __launch_bounds__(kJacThreads) jacobi_eigh_kernel(
    int n, const T* A, long lda, long sA, const double* __restrict__ add, long sadd,
    double* __restrict__ w, long sw, T* Z, long ldz, long sZ, int max_sweeps, double tol,
    int* __restrict__ sweeps) {
  extern __shared__ __align__(16) char smem_raw[];
  JacobiShared<T>& sh = *reinterpret_cast<JacobiShared<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const long b = blockIdx.x;
  const int n2 = n + (n & 1), h = n2 / 2, L = n2;
  const T* Ab = A + b * sA;

  double part = 0.0;
  for (int idx = tid; idx < n2 * n2; idx += kJacThreads) {
    const int i = idx / n2, j = idx - i * n2;
    T v = T(0.0);
    if (i < n && j < n) {
      v = Ab[(long)i * lda + j];
      if (i == j && add) v = v + T(add[b * sadd + i]);
    }
    sh.A[i * L + j] = v;
    sh.V[i * L + j] = T(i == j ? 1.0 : 0.0);
    part += jabs2(v);
  }
  const double nrm2 = jac_block_sum(part, sh.red);
  const double lim2 = tol * tol * nrm2;

  int sweep = 0;
  bool done = (n2 <= 1);
  while (!done && sweep < max_sweeps) {
    if (tid == 0) sh.nrot = 0;
    for (int r = 0; r < n2 - 1; ++r) {
      __syncthreads();
      if (tid < h) {
        // round-robin: position 0 holds index 0, positions 1..n2-1 rotate by one per round
        const int pos_p = tid, pos_q = n2 - 1 - tid;
        const int p = pos_p == 0 ? 0 : 1 + (pos_p - 1 + r) % (n2 - 1);
        const int q = 1 + (pos_q - 1 + r) % (n2 - 1);
        const double a = jre(sh.A[p * L + p]), d = jre(sh.A[q * L + q]);
        const T cpq = sh.A[p * L + q];
        const double ac2 = jabs2(cpq);
        int on = 0;
        double cs = 1.0, sn = 0.0, dpv = a, dqv = d;
        T e = T(1.0);
        if (ac2 > 0.0 && ac2 > 1e-32 * fabs(a) * fabs(d)) {
          const double ac = sqrt(ac2);
          e = jscale(1.0 / ac, jconj(cpq));
          const double tau = (d - a) / (2.0 * ac);
          const double at = fabs(tau);
          double t = at > 1e100 ? 0.5 / at : 1.0 / (at + sqrt(1.0 + tau * tau));
          if (tau < 0.0) t = -t;
          cs = 1.0 / sqrt(1.0 + t * t);
          sn = t * cs;
          dpv = a - t * ac;
          dqv = d + t * ac;
          on = 1;
          atomicAdd(&sh.nrot, 1);
        }
        sh.p[tid] = p;
        sh.q[tid] = q;
        sh.on[tid] = on;
        sh.c[tid] = cs;
        sh.s[tid] = sn;
        sh.e[tid] = e;
        sh.dp[tid] = dpv;
        sh.dq[tid] = dqv;
      }
      __syncthreads();
      // A <- U^H A U, one 2x2 block (rows of pair k, columns of pair l) per item
      for (int it = tid; it < h * h; it += kJacThreads) {
        const int k = it / h, l = it - k * h;
        const int onk = sh.on[k], onl = sh.on[l];
        if (!onk && !onl) continue;
        const int p1 = sh.p[k], q1 = sh.q[k], p2 = sh.p[l], q2 = sh.q[l];
        if (k == l) {
          sh.A[p1 * L + p1] = T(sh.dp[k]);
          sh.A[q1 * L + q1] = T(sh.dq[k]);
          sh.A[p1 * L + q1] = T(0.0);
          sh.A[q1 * L + p1] = T(0.0);
          continue;
        }
        T x00 = sh.A[p1 * L + p2], x01 = sh.A[p1 * L + q2];
        T x10 = sh.A[q1 * L + p2], x11 = sh.A[q1 * L + q2];
        if (onl) {
          const double c2 = sh.c[l], s2 = sh.s[l];
          const T e2 = sh.e[l];
          const T y01 = jmul(x01, e2), y11 = jmul(x11, e2);
          const T n00 = jscale(c2, x00) - jscale(s2, y01), n01 = jscale(s2, x00) + jscale(c2, y01);
          const T n10 = jscale(c2, x10) - jscale(s2, y11), n11 = jscale(s2, x10) + jscale(c2, y11);
          x00 = n00; x01 = n01; x10 = n10; x11 = n11;
        }
        if (onk) {
          const double c1 = sh.c[k], s1 = sh.s[k];
          const T f1 = jconj(sh.e[k]);
          const T y10 = jmul(f1, x10), y11 = jmul(f1, x11);
          const T n00 = jscale(c1, x00) - jscale(s1, y10), n01 = jscale(c1, x01) - jscale(s1, y11);
          const T n10 = jscale(s1, x00) + jscale(c1, y10), n11 = jscale(s1, x01) + jscale(c1, y11);
          x00 = n00; x01 = n01; x10 = n10; x11 = n11;
        }
        sh.A[p1 * L + p2] = x00;
        sh.A[p1 * L + q2] = x01;
        sh.A[q1 * L + p2] = x10;
        sh.A[q1 * L + q2] = x11;
      }
      // V <- V U, one row piece (row i, pair l) per item
      for (int it = tid; it < n2 * h; it += kJacThreads) {
        const int i = it / h, l = it - i * h;
        if (!sh.on[l]) continue;
        const int p2 = sh.p[l], q2 = sh.q[l];
        const double c2 = sh.c[l], s2 = sh.s[l];
        const T x0 = sh.V[i * L + p2], y1 = jmul(sh.V[i * L + q2], sh.e[l]);
        sh.V[i * L + p2] = jscale(c2, x0) - jscale(s2, y1);
        sh.V[i * L + q2] = jscale(s2, x0) + jscale(c2, y1);
      }
    }
    __syncthreads();
    ++sweep;
    double off = 0.0;
    for (int idx = tid; idx < n2 * n2; idx += kJacThreads) {
      const int i = idx / n2, j = idx - i * n2;
      if (i != j) off += jabs2(sh.A[i * L + j]);
    }
    const double off2 = jac_block_sum(off, sh.red);   // ends with a barrier: nrot is final
    done = (sh.nrot == 0) || (off2 <= lim2);
    __syncthreads();   // every thread has read nrot before the next sweep clears it
  }
  __syncthreads();

  // ascending order, ties by index
  if (tid < n) {
    const double wj = jre(sh.A[tid * L + tid]);
    int rk = 0;
    for (int i = 0; i < n; ++i) {
      const double wi = jre(sh.A[i * L + i]);
      rk += (wi < wj) || (wi == wj && i < tid);
    }
    sh.rank[tid] = rk;
    w[b * sw + rk] = wj;
  }
  __syncthreads();
  T* Zb = Z + b * sZ;
  for (int idx = tid; idx < n * n; idx += kJacThreads) {
    const int i = idx / n, j = idx - i * n;
    Zb[(long)i * ldz + sh.rank[j]] = sh.V[i * L + j];
  }
  if (sweeps && tid == 0) sweeps[b] = done ? sweep : -sweep;
}

template <typename T>
int launch_jacobi(long batch, int n, const void* A, long lda, long sA, const double* add,
                  long sadd, double* w, long sw, void* Z, long ldz, long sZ, int max_sweeps,
                  double tol, int* sweeps, hipStream_t st) {
  const size_t smem = sizeof(JacobiShared<T>);   // 133 KB for complex128: one workgroup per CU
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&jacobi_eigh_kernel<T>),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  TNH_REQUIRE(e == hipSuccess, TNH_ERR_HIP, "tnh_eigh_jacobi: %zu bytes of LDS: %s", smem,
              hipGetErrorString(e));
  hipLaunchKernelGGL(jacobi_eigh_kernel<T>, dim3((unsigned)batch), dim3(kJacThreads), smem, st, n,
                     static_cast<const T*>(A), lda, sA, add, sadd, w, sw, static_cast<T*>(Z), ldz,
                     sZ, max_sweeps, tol, sweeps);
  TNH_CHECK_LAUNCH("jacobi_eigh");
  return TNH_OK;
}

}  // namespace tnh

extern "C" int tnh_eigh_jacobi(int dtype, long batch, int n, const void* A, long lda, long sA,
                               const double* add, long sadd, double* w, long sw, void* Z, long ldz,
                               long sZ, int max_sweeps, double tol, int* sweeps,
                               tnh_stream_t stream) {
  TNH_REQUIRE(dtype == TNH_F64 || dtype == TNH_C128, TNH_ERR_UNSUPPORTED,
              "tnh_eigh_jacobi: dtype %d (float64 or complex128 only)", dtype);
  TNH_REQUIRE(n >= 1 && n <= tnh::kJacMax, TNH_ERR_UNSUPPORTED,
              "tnh_eigh_jacobi: n = %d outside 1..%d", n, tnh::kJacMax);
  TNH_REQUIRE(batch >= 0 && batch < (1L << 31) && lda >= n && ldz >= n && sA >= 0 && sZ >= 0 &&
                  sw >= 0 && sadd >= 0 && max_sweeps >= 1 && tol >= 0.0,
              TNH_ERR_ARG, "tnh_eigh_jacobi: bad sizes batch=%ld n=%d lda=%ld ldz=%ld", batch, n,
              lda, ldz);
  if (batch == 0) return TNH_OK;
  TNH_REQUIRE(batch == 1 || (sA >= (long)n * lda - (lda - n) && sZ >= (long)n * ldz - (ldz - n) &&
                             sw >= n && (!add || sadd >= n)),
              TNH_ERR_ARG, "tnh_eigh_jacobi: batch strides overlap (sA=%ld sZ=%ld sw=%ld)", sA, sZ,
              sw);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == TNH_F64)
    return tnh::launch_jacobi<double>(batch, n, A, lda, sA, add, sadd, w, sw, Z, ldz, sZ,
                                      max_sweeps, tol, sweeps, st);
  return tnh::launch_jacobi<tnh::c128>(batch, n, A, lda, sA, add, sadd, w, sw, Z, ldz, sZ,
                                       max_sweeps, tol, sweeps, st);
}
