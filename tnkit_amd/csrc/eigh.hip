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
// Frobenius norm <= tol * ||A||_F. Only the blocks k <= l are transformed; their mirrors are
// written as conjugate transposes.
//
// Used for the Gram matrices of the truncation (tnkit/linalg.py:138-159, the batched form of
// tnkit/batch_tdvp.py:572-599), where rocSOLVER's batched eigh spends milliseconds per call on
// matrices this size.
#include "cplx.h"
#include "tnh_common.h"

namespace tnh {

constexpr int kJacMax = 64;
constexpr int kEighThreads = 512;   // the eigh kernel: eight waves share a round's block updates

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
  unsigned short kl[kJacMax / 2 * (kJacMax / 2 + 1) / 2];   // block pairs k <= l, (k << 8) | l
  int rank[kJacMax];
  double red[kEighThreads / kWave];
  int nrot;
};

template <int NT>
__device__ __forceinline__ double jac_block_sum(double v, double* red) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  __syncthreads();
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = v;
  __syncthreads();
  double t = 0.0;
#pragma unroll
  for (int w = 0; w < NT / kWave; ++w) t += red[w];
  return t;
}

// matrix b: A_b = A + b*sA (row i at i*lda), plus add_b[j] on the diagonal when add is set;
// w_b ascending eigenvalues, Z_b[i*ldz + j] = component i of eigenvector j (Z may alias A).
// sweeps[b] (optional) = sweeps run, negated when max_sweeps ended the solve unconverged.
template <typename T>
__global__ void __launch_bounds__(kEighThreads) jacobi_eigh_kernel(
    int n, const T* A, long lda, long sA, const double* __restrict__ add, long sadd,
    double* __restrict__ w, long sw, T* Z, long ldz, long sZ, int max_sweeps, double tol,
    int* __restrict__ sweeps) {
  extern __shared__ __align__(16) char smem_raw[];
  JacobiShared<T>& sh = *reinterpret_cast<JacobiShared<T>*>(smem_raw);
  const int tid = threadIdx.x;
  const long b = blockIdx.x;
  const int n2 = n + (n & 1), h = n2 / 2, L = n2;
  const T* Ab = A + b * sA;

  // the Hermitian part of A (both triangles read, the diagonal real), padded with zeros
  double part = 0.0;
  for (int idx = tid; idx < n2 * n2; idx += kEighThreads) {
    const int i = idx / n2, j = idx - i * n2;
    T v = T(0.0);
    if (i < n && j < n) {
      if (i == j) {
        v = T(jre(Ab[(long)i * lda + i]) + (add ? add[b * sadd + i] : 0.0));
      } else {
        v = jscale(0.5, Ab[(long)i * lda + j] + jconj(Ab[(long)j * lda + i]));
      }
    }
    sh.A[i * L + j] = v;
    sh.V[i * L + j] = T(i == j ? 1.0 : 0.0);
    part += jabs2(v);
  }
  // the blocks of one round: pairs k <= l of the h index pairs (the others are their mirrors)
  const int nkl = h * (h + 1) / 2;
  for (int it = tid; it < nkl; it += kEighThreads) {
    int k = 0, rem = it;
    while (rem >= h - k) {
      rem -= h - k;
      ++k;
    }
    sh.kl[it] = static_cast<unsigned short>((k << 8) | (k + rem));
  }
  const double nrm2 = jac_block_sum<kEighThreads>(part, sh.red);
  const double lim2 = tol * tol * nrm2;

  int sweep = 0;
  // a non-finite matrix is reported unconverged at once: its NaN entries would never rotate and
  // the sweeps would end "converged" on a partly diagonalised matrix
  bool bad = !isfinite(nrm2);
  bool done = (n2 <= 1);
  while (!done && !bad && sweep < max_sweeps) {
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
          // few long-latency operations on the critical path: one rsqrt for |c| and the phase,
          // one sqrt and one division for t, one rsqrt for the cosine
          const double inv = rsqrt(ac2);
          const double ac = ac2 * inv;
          e = jscale(inv, jconj(cpq));
          const double tau = 0.5 * (d - a) * inv;
          const double at = fabs(tau);
          double t = at > 1e100 ? 0.5 / at : 1.0 / (at + sqrt(1.0 + tau * tau));
          if (tau < 0.0) t = -t;
          cs = rsqrt(1.0 + t * t);
          sn = t * cs;
          dpv = a - t * ac;
          dqv = d + t * ac;
          on = 1;
          sh.nrot = 1;   // any rotation this sweep (every writer stores the same value)
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
      // A <- U^H A U, one 2x2 block (rows of pair k, columns of pair l, k <= l) per item; the
      // block (l, k) is its conjugate transpose, so A stays exactly Hermitian
      for (int it = tid; it < nkl; it += kEighThreads) {
        const int k = sh.kl[it] >> 8, l = sh.kl[it] & 255;
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
        sh.A[p2 * L + p1] = jconj(x00);
        sh.A[q2 * L + p1] = jconj(x01);
        sh.A[p2 * L + q1] = jconj(x10);
        sh.A[q2 * L + q1] = jconj(x11);
      }
      // V <- V U, one row piece (row i, pair l) per item
      for (int it = tid; it < n2 * h; it += kEighThreads) {
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
    for (int idx = tid; idx < n2 * n2; idx += kEighThreads) {
      const int i = idx / n2, j = idx - i * n2;
      if (i != j) off += jabs2(sh.A[i * L + j]);
    }
    const double off2 = jac_block_sum<kEighThreads>(off, sh.red);   // ends with a barrier: nrot is final
    bad = !isfinite(off2);
    done = !bad && ((sh.nrot == 0) || (off2 <= lim2));
    __syncthreads();   // every thread has read nrot before the next sweep clears it
  }
  __syncthreads();

  // ascending order, ties by index; a NaN sorts last (a total order, so the ranks stay a
  // permutation and every output is written; the caller sees the NaN)
  if (tid < n) {
    const double wj = jre(sh.A[tid * L + tid]);
    const double kj = isnan(wj) ? INFINITY : wj;
    int rk = 0;
    for (int i = 0; i < n; ++i) {
      const double wi = jre(sh.A[i * L + i]);
      const double ki = isnan(wi) ? INFINITY : wi;
      rk += (ki < kj) || (ki == kj && i < tid);
    }
    sh.rank[tid] = rk;
    w[b * sw + rk] = wj;
  }
  __syncthreads();
  T* Zb = Z + b * sZ;
  for (int idx = tid; idx < n * n; idx += kEighThreads) {
    const int i = idx / n, j = idx - i * n;
    Zb[(long)i * ldz + sh.rank[j]] = sh.V[i * L + j];
  }
  if (sweeps && tid == 0) sweeps[b] = (done && !bad) ? sweep : -(sweep > 0 ? sweep : 1);
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
  hipLaunchKernelGGL(jacobi_eigh_kernel<T>, dim3((unsigned)batch), dim3(kEighThreads), smem, st, n,
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

namespace tnh {

// ---- thin SVD of small matrices: one-sided (Hestenes) Jacobi -----------------------------------
// For A (m x n), the work matrix B = A (m >= n) or A^H (m < n), M x N with N = min(m, n) <= 64,
// sits in LDS column-major with the accumulated right rotations V (N x N). A sweep visits every
// column pair once in the round-robin rounds of jacobi_eigh_kernel; a group of G lanes (a power
// of two within one wave) owns one pair of a round: it forms |b_p|^2, |b_q|^2 and b_p^H b_q,
// and when |b_p^H b_q| > tol |b_p| |b_q| rotates both columns of B and of V by the Jacobi
// rotation of that 2 x 2 Gram matrix (orthogonalising them). Pairs of a round are disjoint, so a
// round needs one barrier. At convergence (a sweep without rotations) the column norms are the
// singular values, the normalised columns the left vectors of B, V its right vectors; columns
// are orthogonal to tol relative to their own norms, so small singular values keep their
// relative accuracy. Exactly zero columns (rank deficiency) get an orthonormal completion by
// two Gram-Schmidt passes over unit vectors, so the returned U and Vh are always isometries.
template <typename T>
__device__ __forceinline__ T jshfl_xor(T v, int off);
template <>
__device__ __forceinline__ double jshfl_xor<double>(double v, int off) {
  return __shfl_xor(v, off, 64);
}
template <>
__device__ __forceinline__ c128 jshfl_xor<c128>(c128 v, int off) {
  return {__shfl_xor(v.re, off, 64), __shfl_xor(v.im, off, 64)};
}

// LDS of one matrix: W (M x N2), V (N2 x N2) and N2 coefficients in T; N2 singular values,
// ranks and flags; 64 doubles of reduction scratch
__host__ __device__ constexpr size_t svd_jacobi_lds(int M, int N2, size_t elem) {
  return ((size_t)M * N2 + (size_t)N2 * N2 + N2) * elem +
         (size_t)N2 * (sizeof(double) + 2 * sizeof(int)) + 64 * sizeof(double);
}
constexpr size_t kSvdLdsMax = 163840 - 1024;   // 160 KB of LDS, less the static variables

template <typename T, int NT>
__global__ void __launch_bounds__(NT) svd_jacobi_kernel(
    int m, int n, const T* __restrict__ A, long lda, long sA, double* __restrict__ S, long sS,
    T* __restrict__ U, long ldu, long sU, T* __restrict__ Vh, long ldv, long sV, int max_sweeps,
    double tol, int* __restrict__ sweeps) {
  extern __shared__ __align__(16) char smem_raw[];
  const bool tr = m < n;
  const int M = tr ? n : m, N = tr ? m : n, N2 = N + (N & 1), h = N2 / 2;
  T* W = reinterpret_cast<T*>(smem_raw);               // M x N2, column j at W + j*M
  T* V = W + (size_t)M * N2;                           // N2 x N2, column j at V + j*N2
  T* coef = V + (size_t)N2 * N2;                       // N2 projection coefficients
  double* sig = reinterpret_cast<double*>(coef + N2);
  int* rank = reinterpret_cast<int*>(sig + N2);
  int* valid = rank + N2;
  double* red = reinterpret_cast<double*>(valid + N2);  // 64 doubles of scratch
  __shared__ int nrot;
  const int tid = threadIdx.x;
  const long b = blockIdx.x;
  const T* Ab = A + b * sA;

  for (int idx = tid; idx < M * N2; idx += NT) {
    const int j = idx / M, i = idx - j * M;
    T v = T(0.0);
    if (j < N) v = tr ? jconj(Ab[(long)j * lda + i]) : Ab[(long)i * lda + j];
    W[idx] = v;
  }
  for (int idx = tid; idx < N2 * N2; idx += NT) {
    const int j = idx / N2, i = idx - j * N2;
    V[idx] = T(i == j ? 1.0 : 0.0);
  }
  // lanes per pair: the largest power of two <= min(64, 256 / h)
  int G = 64;
  while (G * h > NT) G >>= 1;
  const int g = tid / G, gl = tid - g * G;
  const bool active = g < h;

  int sweep = 0;
  bool done = false;
  while (!done && sweep < max_sweeps) {
    if (tid == 0) nrot = 0;
    for (int r = 0; r < N2 - 1; ++r) {
      __syncthreads();
      if (active) {
        const int pos_q = N2 - 1 - g;
        const int p = g == 0 ? 0 : 1 + (g - 1 + r) % (N2 - 1);
        const int q = 1 + (pos_q - 1 + r) % (N2 - 1);
        T* wp = W + (size_t)p * M;
        T* wq = W + (size_t)q * M;
        double a = 0.0, d = 0.0;
        T gm = T(0.0);
        for (int i = gl; i < M; i += G) {
          const T x = wp[i], y = wq[i];
          a += jabs2(x);
          d += jabs2(y);
          gm = gm + jmul(jconj(x), y);
        }
        for (int off = G >> 1; off >= 1; off >>= 1) {
          a += __shfl_xor(a, off, 64);
          d += __shfl_xor(d, off, 64);
          gm = gm + jshfl_xor(gm, off);
        }
        const double ag2 = jabs2(gm);
        if (ag2 > 0.0 && ag2 > tol * tol * a * d) {
          const double inv = rsqrt(ag2);
          const T e = jscale(inv, jconj(gm));
          const double tau = 0.5 * (d - a) * inv;
          const double at = fabs(tau);
          double t = at > 1e100 ? 0.5 / at : 1.0 / (at + sqrt(1.0 + tau * tau));
          if (tau < 0.0) t = -t;
          const double cs = rsqrt(1.0 + t * t), sn = t * cs;
          for (int i = gl; i < M; i += G) {
            const T x = wp[i], y = jmul(wq[i], e);
            wp[i] = jscale(cs, x) - jscale(sn, y);
            wq[i] = jscale(sn, x) + jscale(cs, y);
          }
          T* vp = V + (size_t)p * N2;
          T* vq = V + (size_t)q * N2;
          for (int i = gl; i < N2; i += G) {
            const T x = vp[i], y = jmul(vq[i], e);
            vp[i] = jscale(cs, x) - jscale(sn, y);
            vq[i] = jscale(sn, x) + jscale(cs, y);
          }
          if (gl == 0) nrot = 1;   // any rotation this sweep (same value from every writer)
        }
      }
    }
    __syncthreads();
    ++sweep;
    done = (nrot == 0);
    __syncthreads();   // every thread has read nrot before the next sweep clears it
  }

  // singular values and descending order (ties by index)
  for (int j = tid; j < N; j += NT) {
    double s2 = 0.0;
    for (int i = 0; i < M; ++i) s2 += jabs2(W[(size_t)j * M + i]);
    sig[j] = sqrt(s2);
  }
  __syncthreads();
  for (int j = tid; j < N; j += NT) {
    const double sj = sig[j];
    const double kj = isnan(sj) ? -1.0 : sj;   // NaN sorts last (ranks stay a permutation)
    int rk = 0;
    for (int i = 0; i < N; ++i) {
      const double ki = isnan(sig[i]) ? -1.0 : sig[i];
      rk += (ki > kj) || (ki == kj && i < j);
    }
    rank[j] = rk;
    valid[j] = sj > 0.0;
  }
  __syncthreads();
  for (int idx = tid; idx < M * N; idx += NT) {
    const int j = idx / M;
    if (sig[j] > 0.0) W[idx] = jscale(1.0 / sig[j], W[idx]);
  }
  __syncthreads();
  // orthonormal completion of exactly zero columns (uniform loops: sig and valid are in LDS)
  for (int j = 0; j < N; ++j) {
    if (sig[j] > 0.0) continue;
    for (int cand = 0; cand < M; ++cand) {
      // column j <- e_cand, then two passes of  x <- x - sum_k w_k (w_k^H x)  over valid k
      for (int i = tid; i < M; i += NT) W[(size_t)j * M + i] = T(i == cand ? 1.0 : 0.0);
      __syncthreads();
      for (int pass = 0; pass < 2; ++pass) {
        // coefficients c_k = w_k^H x over the valid columns, one thread per k
        for (int k = tid; k < N; k += NT) {
          T c = T(0.0);
          if (valid[k])
            for (int i = 0; i < M; ++i)
              c = c + jmul(jconj(W[(size_t)k * M + i]), W[(size_t)j * M + i]);
          coef[k] = c;
        }
        __syncthreads();
        for (int i = tid; i < M; i += NT) {
          T x = W[(size_t)j * M + i];
          for (int k = 0; k < N; ++k)
            if (valid[k]) x = x - jmul(W[(size_t)k * M + i], coef[k]);
          W[(size_t)j * M + i] = x;
        }
        __syncthreads();
      }
      double part = 0.0;
      for (int i = tid; i < M; i += NT) part += jabs2(W[(size_t)j * M + i]);
      const double nrm2 = jac_block_sum<NT>(part, red);
      if (nrm2 > 0.5 / M) {   // some e_cand keeps >= (M - rank) / M >= 1 / M of its norm
        const double inv = 1.0 / sqrt(nrm2);
        for (int i = tid; i < M; i += NT)
          W[(size_t)j * M + i] = jscale(inv, W[(size_t)j * M + i]);
        __syncthreads();
        if (tid == 0) valid[j] = 1;
        __syncthreads();
        break;
      }
      __syncthreads();
    }
  }
  __syncthreads();

  // outputs: k = N singular triplets, descending
  double* Sb = S + b * sS;
  T* Ub = U + b * sU;
  T* Vb = Vh + b * sV;
  for (int j = tid; j < N; j += NT) Sb[rank[j]] = sig[j];
  if (!tr) {
    // U (m x N): column rank[j] = w_j; Vh (N x n): row rank[j] = v_j^H
    for (int idx = tid; idx < M * N; idx += NT) {
      const int j = idx / M, i = idx - j * M;
      Ub[(long)i * ldu + rank[j]] = W[idx];
    }
    for (int idx = tid; idx < N * N; idx += NT) {
      const int j = idx / N, c = idx - j * N;
      Vb[(long)rank[j] * ldv + c] = jconj(V[(size_t)j * N2 + c]);
    }
  } else {
    // A = V S W^H: U (m = N x N): column rank[j] = v_j; Vh (N x n = M): row rank[j] = w_j^H
    for (int idx = tid; idx < N * N; idx += NT) {
      const int j = idx / N, i = idx - j * N;
      Ub[(long)i * ldu + rank[j]] = V[(size_t)j * N2 + i];
    }
    for (int idx = tid; idx < M * N; idx += NT) {
      const int j = idx / M, c = idx - j * M;
      Vb[(long)rank[j] * ldv + c] = jconj(W[idx]);
    }
  }
  if (sweeps && tid == 0) {   // non-finite input shows up as a non-finite singular value
    bool fin = true;
    for (int j = 0; j < N; ++j) fin = fin && isfinite(sig[j]);
    sweeps[b] = (done && fin) ? sweep : -(sweep > 0 ? sweep : 1);
  }
}

template <typename T>
int launch_svd_jacobi(long batch, int m, int n, const void* A, long lda, long sA, double* S,
                      long sS, void* U, long ldu, long sU, void* Vh, long ldv, long sV,
                      int max_sweeps, double tol, int* sweeps, hipStream_t st) {
  const int M = m > n ? m : n, N = m < n ? m : n, N2 = N + (N & 1);
  const size_t smem = svd_jacobi_lds(M, N2, sizeof(T));
  // eight waves (16 lanes per column pair) pay off from N > 32; below, four waves
  auto kern = N > 32 ? &svd_jacobi_kernel<T, 512> : &svd_jacobi_kernel<T, 256>;
  const int nt = N > 32 ? 512 : 256;
  hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(kern),
                                     hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
  TNH_REQUIRE(e == hipSuccess, TNH_ERR_HIP, "tnh_svd_jacobi: %zu bytes of LDS: %s", smem,
              hipGetErrorString(e));
  hipLaunchKernelGGL(kern, dim3((unsigned)batch), dim3(nt), smem, st, m, n,
                     static_cast<const T*>(A), lda, sA, S, sS, static_cast<T*>(U), ldu, sU,
                     static_cast<T*>(Vh), ldv, sV, max_sweeps, tol, sweeps);
  TNH_CHECK_LAUNCH("svd_jacobi");
  return TNH_OK;
}

}  // namespace tnh

extern "C" size_t tnh_svd_jacobi_lds_bytes(int dtype, int m, int n) {
  const int M = m > n ? m : n, N = m < n ? m : n;
  const size_t elem = dtype == TNH_C128 ? 16 : 8;
  return tnh::svd_jacobi_lds(M, N + (N & 1), elem);
}

extern "C" int tnh_svd_jacobi(int dtype, long batch, int m, int n, const void* A, long lda,
                              long sA, double* S, long sS, void* U, long ldu, long sU, void* Vh,
                              long ldv, long sV, int max_sweeps, double tol, int* sweeps,
                              tnh_stream_t stream) {
  TNH_REQUIRE(dtype == TNH_F64 || dtype == TNH_C128, TNH_ERR_UNSUPPORTED,
              "tnh_svd_jacobi: dtype %d (float64 or complex128 only)", dtype);
  const int N = m < n ? m : n;
  TNH_REQUIRE(m >= 1 && n >= 1 && N <= tnh::kJacMax, TNH_ERR_UNSUPPORTED,
              "tnh_svd_jacobi: min(m, n) = %d outside 1..%d", N, tnh::kJacMax);
  TNH_REQUIRE(tnh_svd_jacobi_lds_bytes(dtype, m, n) <= tnh::kSvdLdsMax, TNH_ERR_UNSUPPORTED,
              "tnh_svd_jacobi: %d x %d needs %zu bytes of LDS (at most %zu)", m, n,
              tnh_svd_jacobi_lds_bytes(dtype, m, n), tnh::kSvdLdsMax);
  TNH_REQUIRE(batch >= 0 && batch < (1L << 31) && lda >= n && ldu >= N && ldv >= n && sA >= 0 &&
                  sS >= 0 && sU >= 0 && sV >= 0 && max_sweeps >= 1 && tol > 0.0,
              TNH_ERR_ARG, "tnh_svd_jacobi: bad sizes batch=%ld m=%d n=%d lda=%ld ldu=%ld ldv=%ld",
              batch, m, n, lda, ldu, ldv);
  if (batch == 0) return TNH_OK;
  // each matrix's outputs must not overlap the next one's (A is only read, so it may repeat)
  TNH_REQUIRE(batch == 1 || (sU >= (long)(m - 1) * ldu + N && sV >= (long)(N - 1) * ldv + n &&
                             sS >= N),
              TNH_ERR_ARG, "tnh_svd_jacobi: output batch strides overlap (sU=%ld sV=%ld sS=%ld)",
              sU, sV, sS);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == TNH_F64)
    return tnh::launch_svd_jacobi<double>(batch, m, n, A, lda, sA, S, sS, U, ldu, sU, Vh, ldv, sV,
                                          max_sweeps, tol, sweeps, st);
  return tnh::launch_svd_jacobi<tnh::c128>(batch, m, n, A, lda, sA, S, sS, U, ldu, sU, Vh, ldv,
                                           sV, max_sweeps, tol, sweeps, st);
}
