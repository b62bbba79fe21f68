// Real symmetric eigensolver for orders up to 2048 (float64): the eigendecomposition behind the
// Gram truncation (tnkit/linalg.py:138-159) at the sweep's wide bonds, where the Gram matrix of
// theta (d chi x d chi) reaches order 2048 at chi = 1024.
//
// Three phases:
//  1. Householder tridiagonalisation A = Q T Q^T (lower form, Q = H_0 H_1 ... H_{n-2},
//     H_j = I - tau_j v_j v_j^T, v_j zero at and above j, v_j[j+1] = 1) in ONE cooperative launch
//     (sytrd_coop_kernel). Row i of the matrix lives in the LDS of workgroup i mod G for the
//     whole reduction; a column costs one grid barrier. After the barrier of column j-1 every
//     workgroup (a) forms w = p - gamma v from the published p = tau A v, (b) applies the rank-2
//     update A -= v w^T + w v^T to its own rows, (c) rebuilds row j from its pre-update copy
//     (published by the row's owner before the barrier) with the same rank-2 correction, so
//     the next reflector is known to every workgroup without a second barrier, then (d) forms
//     its rows of p for the new reflector and publishes them with its partial v.p. The matrix is
//     read from LDS only; what crosses workgroups per column is two vectors of n and one partial
//     per workgroup (double-buffered by column parity).
//  2. Divide and conquer on T (Cuppen's tearing with Gu-Eisenstat eigenvectors), leaves of
//     order <= 32 on the batched Jacobi kernel.
//  3. The wanted eigenvectors of A = Q times those of T, applied as WY blocks of 32 reflectors
//     (two MFMA GEMMs per block).
#include <algorithm>
#include <vector>

#include "cplx.h"
#include "tnh_common.h"

namespace tnh {

// ---- grid barrier ---------------------------------------------------------------------------
// Workgroup g writes the epoch to its own arrival word; workgroup 0 polls all arrival words (one
// per thread) and then writes the epoch to every workgroup's own release word, on which that
// workgroup waits. No word is written or polled by more than one workgroup besides the master,
// so nothing contends (an atomic counter polled by 256 workgroups costs ~11 us per barrier on
// this part; this form ~2.5 us; polling the exchanged data itself instead of a barrier floods
// memory and measured slower). Data that crosses workgroups is written with relaxed agent-scope
// atomics (through to memory) into slots that are written once per launch and read after the
// barrier with ordinary loads (never cached before, so never stale); every wave waits for its
// stores before arriving. Waits are bounded: a wait that runs past the spin limit sets *err (the
// master still releases), and every workgroup leaves at its next barrier, so a broken launch
// still ends.
constexpr int kBarStride = 16;          // unsigned words between barrier words (64 bytes)
constexpr int kBarSpin = 1 << 20;
constexpr int kBarMaxG = 256;
constexpr size_t kBarWords = (size_t)2 * kBarMaxG * kBarStride;

struct GridBar {
  unsigned* mem;    // arrival word of g at g * kBarStride, release word at (kBarMaxG + g) * stride
  int* err;
};

__device__ __forceinline__ void st_shared(double* p, double v) {
  __hip_atomic_store(p, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// The barrier over the G workgroups g = 0..G-1 of one group (a grid can hold several groups,
// each with its own barrier words).
__device__ bool grid_sync_of(const GridBar& b, unsigned epoch, int* flag, unsigned G, unsigned g) {
  __builtin_amdgcn_s_waitcnt(0);   // this wave's stores are complete
  __syncthreads();
  if (threadIdx.x == 0)
    __hip_atomic_store(b.mem + g * kBarStride, epoch, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  if (g == 0) {
    int bad = 0;
    for (unsigned q = threadIdx.x; q < G; q += blockDim.x) {
      int spins = 0;
      while (__hip_atomic_load(b.mem + q * kBarStride, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT) < epoch) {
        if (++spins > kBarSpin) {
          bad = 1;
          break;
        }
        __builtin_amdgcn_s_sleep(1);
      }
    }
    if (bad) __hip_atomic_store(b.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    __builtin_amdgcn_s_waitcnt(0);
    __syncthreads();
    for (unsigned q = threadIdx.x; q < G; q += blockDim.x)
      __hip_atomic_store(b.mem + (kBarMaxG + q) * kBarStride, epoch, __ATOMIC_RELAXED,
                         __HIP_MEMORY_SCOPE_AGENT);
  }
  if (threadIdx.x == 0) {
    int spins = 0, bad = 0;
    while (__hip_atomic_load(b.mem + (kBarMaxG + g) * kBarStride, __ATOMIC_RELAXED,
                             __HIP_MEMORY_SCOPE_AGENT) < epoch) {
      if (++spins > kBarSpin) {
        bad = 1;
        break;
      }
      __builtin_amdgcn_s_sleep(1);
    }
    if (bad) __hip_atomic_store(b.err, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    *flag = bad | __hip_atomic_load(b.err, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
  }
  __syncthreads();
  return *flag == 0;
}

__device__ bool grid_sync(const GridBar& b, unsigned epoch, int* flag) {
  return grid_sync_of(b, epoch, flag, gridDim.x, blockIdx.x);
}

__global__ void bar_reset_kernel(unsigned* mem, int* err) {
  for (size_t i = threadIdx.x; i < kBarWords; i += blockDim.x) mem[i] = 0u;
  if (threadIdx.x == 0) *err = 0;
}

// ---- phase 1: tridiagonalisation -------------------------------------------------------------
constexpr int kTrdRows = 8;                     // rows per workgroup
constexpr int kSyevMaxN = 2048;                 // threads x columns per thread at the largest

// Wave sum on VALU only: four DPP lane permutations (xor 1 and xor 2 inside quads, half-row and
// row mirrors) leave every lane with its 16-lane row's sum, and the four row sums are read out of
// lanes 0, 16, 32 and 48 (readlane) and added in a fixed order, so the total is the same
// (bitwise) in every lane and every workgroup.
template <int CTRL>
__device__ __forceinline__ double dpp_d(double v) {
  const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, 0xF, 0xF, false);
  const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, 0xF, 0xF, false);
  return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double lane_d(double v, int l) {
  return __hiloint2double(__builtin_amdgcn_readlane(__double2hiint(v), l),
                          __builtin_amdgcn_readlane(__double2loint(v), l));
}

__device__ __forceinline__ double wsum(double v) {
  v += dpp_d<0xB1>(v);    // quad_perm [1, 0, 3, 2]
  v += dpp_d<0x4E>(v);    // quad_perm [2, 3, 0, 1]
  v += dpp_d<0x141>(v);   // row_half_mirror
  v += dpp_d<0x140>(v);   // row_mirror
  return (lane_d(v, 0) + lane_d(v, 16)) + (lane_d(v, 32) + lane_d(v, 48));
}

// Wave sums of eight per-lane values s[0..7] at once (the eight own rows of a workgroup's pass):
// two halving exchanges (xor 1, then xor 2) leave each lane two of the eight rows, summed over
// its lane pair and quad, then row rotations by 4 and 8 and exchanges across the four 16-lane
// rows finish the sums. Lanes 0..3 hold rows 4 b0 + 2 b1 + {0, 1} (b0, b1 the lane's low bits)
// and write them to dst[0..7]. About a quarter of the instructions of eight wsum calls.
__device__ __forceinline__ void wsum8_store(const double (&s)[8], double* dst, int lane) {
  const bool b0 = lane & 1, b1 = (lane >> 1) & 1;
  double t[4];
#pragma unroll
  for (int i = 0; i < 4; ++i) {
    const double keep = b0 ? s[i + 4] : s[i];
    const double send = b0 ? s[i] : s[i + 4];
    t[i] = keep + dpp_d<0xB1>(send);      // quad_perm [1, 0, 3, 2]: lane ^ 1
  }
  double u[2];
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    const double keep = b1 ? t[i + 2] : t[i];
    const double send = b1 ? t[i] : t[i + 2];
    u[i] = keep + dpp_d<0x4E>(send);      // quad_perm [2, 3, 0, 1]: lane ^ 2
  }
#pragma unroll
  for (int i = 0; i < 2; ++i) {
    u[i] += dpp_d<0x124>(u[i]);           // row_ror:4
    u[i] += dpp_d<0x128>(u[i]);           // row_ror:8
    u[i] += __shfl_xor(u[i], 16, 64);
    u[i] += __shfl_xor(u[i], 32, 64);
  }
  if (lane < 4) {
    dst[4 * b0 + 2 * b1] = u[0];
    dst[4 * b0 + 2 * b1 + 1] = u[1];
  }
}

struct TrdArgs {
  int n;
  int ns;         // slot length of P and Rp (n rounded up, plus a gap)
  int gs;         // slot length of part
  const double* A;
  long lda;
  double* Vr;     // n x n, row j = v_j
  double* d;      // n
  double* e;      // n - 1
  double* tau;    // n - 1
  double* P;      // slot j: tau_j A v_j (rows > j)
  double* Rp;     // slot j: row j before its update by (v_{j-1}, w_{j-1}) (columns >= j)
  double* part;   // slot j: per-workgroup partial v_j . p_j
  GridBar bar;
};

// Sum over the workgroup of one value per thread with a single barrier: red is used by
// alternate calls in two halves, so a call's writes cannot meet the previous call's reads.
template <int kTrdWaves>
__device__ __forceinline__ double trd_sum1(double v, double* red, int& par) {
  v = wsum(v);
  double* rb = red + kTrdWaves * par;
  par ^= 1;
  if ((threadIdx.x & 63) == 0) rb[threadIdx.x >> 6] = v;
  __syncthreads();
  double s = 0.0;
#pragma unroll
  for (int w = 0; w < kTrdWaves; ++w) s += rb[w];
  return s;
}

// Grid of G workgroups (G <= n, ceil(n / G) <= kTrdRows) of kTrdThreads threads, launched
// cooperatively. Row i of A lives in the REGISTERS of workgroup i mod G for the whole reduction:
// thread t holds columns t + kTrdThreads c (c < kTrdCols) of its workgroup's (at most 8) rows,
// so the rank-2 update and the product A v are register arithmetic and the only on-chip traffic
// is the reductions.
// A is read once (row i by workgroup i mod G, as given: the caller passes a symmetric matrix).
//
// Column j after the barrier: gamma from the partials, w_{j-1} = p - gamma v (v_{j-1}[j] = 1),
// row j = its published copy - v w^T - w v^T at row j, the reflector v_j, then one pass over
// the own rows that applies the rank-2 update of column j-1 and accumulates p_j = tau_j A v_j;
// the owner of row j+1 publishes it from the same registers.
template <int kTrdThreads, int kTrdCols>
__global__ void __launch_bounds__(kTrdThreads) sytrd_coop_kernel(TrdArgs a) {
  constexpr int kTrdWaves = kTrdThreads / 64;
  __shared__ double red[2 * kTrdWaves];
  __shared__ double red8[2][kTrdWaves][kTrdRows];
  __shared__ double own_v[kTrdRows], own_p[kTrdRows], own_vn[kTrdRows];
  __shared__ double vnext;    // v_{j-1}[j+1], for row j's subdiagonal entry
  __shared__ int flag;
  const int n = a.n, G = gridDim.x, g = blockIdx.x, t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int nr = (n - g + G - 1) / G;
  int par = 0, par8 = 0;

  double x[kTrdRows][kTrdCols];   // the own rows at this thread's columns
#pragma unroll
  for (int r = 0; r < kTrdRows; ++r)
#pragma unroll
    for (int c = 0; c < kTrdCols; ++c) {
      const int l = t + kTrdThreads * c;
      x[r][c] = (r < nr && l < n) ? a.A[(long)(g + r * G) * a.lda + l] : 0.0;
    }
  if (n == 1) {
    if (t == 0) a.d[0] = x[0][0];
    return;
  }
  if (g == 0)
#pragma unroll
    for (int c = 0; c < kTrdCols; ++c) {
      const int l = t + kTrdThreads * c;
      if (l < n) st_shared(a.Rp + l, x[0][c]);
    }
  if (t < kTrdRows) {
    own_v[t] = 0.0;
    own_p[t] = 0.0;
  }
  unsigned epoch = 1;
  if (!grid_sync(a.bar, epoch++, &flag)) return;

  double vr[kTrdCols];   // v_{j-1} at this thread's columns
#pragma unroll
  for (int c = 0; c < kTrdCols; ++c) vr[c] = 0.0;
  double tau_prev = 0.0;

  for (int j = 0; j < n - 1; ++j) {
    const double* Rj = a.Rp + (long)j * a.ns;
    double R[kTrdCols], wr[kTrdCols];
    double gamma = 0.0, dj, alpha;
    if (j > 0) {
      const double* Pj = a.P + (long)(j - 1) * a.ns;
      const double* pj = a.part + (long)(j - 1) * a.gs;
      // every wave sums the G partials itself, in one fixed order (gamma is then the same in
      // every wave and workgroup without an LDS round trip)
      double ps = 0.0;
#pragma unroll
      for (int q = 0; q < kBarMaxG / 64; ++q) {
        const int k = lane + 64 * q;
        ps += k < G ? pj[k] : 0.0;
      }
      double Pl[kTrdCols];
#pragma unroll
      for (int c = 0; c < kTrdCols; ++c) {
        const int l = t + kTrdThreads * c;
        Pl[c] = l >= j && l < n ? Pj[l] : 0.0;
        R[c] = l >= j && l < n ? Rj[l] : 0.0;
      }
      const double pjj = Pj[j], pj1 = Pj[j + 1], rjj = Rj[j], rj1 = Rj[j + 1];
      gamma = 0.5 * tau_prev * wsum(ps);
      const double wj = pjj - gamma, vx = vnext;
#pragma unroll
      for (int c = 0; c < kTrdCols; ++c) {
        wr[c] = Pl[c] - gamma * vr[c];
        R[c] -= wr[c] + wj * vr[c];
      }
      dj = rjj - 2.0 * wj;
      alpha = rj1 - vx * wj - (pj1 - gamma * vx);
    } else {
#pragma unroll
      for (int c = 0; c < kTrdCols; ++c) {
        const int l = t + kTrdThreads * c;
        R[c] = l < n ? Rj[l] : 0.0;
        wr[c] = 0.0;
      }
      dj = Rj[0];
      alpha = Rj[1];
    }

    // the reflector of column j: x = row j right of the diagonal
    double sq = 0.0;
#pragma unroll
    for (int c = 0; c < kTrdCols; ++c) {
      const int l = t + kTrdThreads * c;
      if (l >= j + 2 && l < n) sq += R[c] * R[c];
    }
    const double sigma = trd_sum1<kTrdWaves>(sq, red, par);
    double beta = alpha, tj = 0.0, scale = 0.0;
    if (sigma > 0.0) {
      beta = -copysign(sqrt(alpha * alpha + sigma), alpha);
      tj = (beta - alpha) / beta;
      scale = 1.0 / (alpha - beta);
    }
    double vn[kTrdCols];
#pragma unroll
    for (int c = 0; c < kTrdCols; ++c) {
      const int l = t + kTrdThreads * c;
      vn[c] = l == j + 1 ? 1.0 : (l >= j + 2 && l < n ? R[c] * scale : 0.0);
      if (l < n && l >= g && (l - g) % G == 0) {
        own_vn[(l - g) / G] = vn[c];
        a.Vr[(long)j * n + l] = vn[c];     // v_j at the own rows; row j of Vr is shared out
      }
      if (l == j + 2) vnext = vn[c];
    }
    if (g == j % G && t == 0) {
      a.d[j] = dj;
      a.e[j] = beta;
      a.tau[j] = tj;
    }
    if (j == n - 2) {   // x has one entry (tau = 0); the last diagonal gets column j-1's update
      if (g == (n - 1) % G) {
        const int rl = (n - 1 - g) / G;
        const double vi = own_v[rl], wi = own_p[rl] - gamma * own_v[rl];
#pragma unroll
        for (int r = 0; r < kTrdRows; ++r)
#pragma unroll
          for (int c = 0; c < kTrdCols; ++c)
            if (r == rl && t + kTrdThreads * c == n - 1) a.d[n - 1] = x[r][c] - 2.0 * vi * wi;
      }
      break;
    }
    // (own_vn and vnext, written just above, are read only after the next barrier of this
    // workgroup and in the next column; the pass reads own_v and own_p, written in the last one)

    // one pass over the own rows: update by (v_{j-1}, w_{j-1}) on columns >= j, p_j on rows > j,
    // row j+1 published by its owner
    const int rpub = (j + 1 < n - 1 && g == (j + 1) % G) ? (j + 1 - g) / G : -1;
    double (*rb)[kTrdRows] = red8[par8];
    par8 ^= 1;
    double sacc[kTrdRows];
#pragma unroll
    for (int r = 0; r < kTrdRows; ++r) {
      const double vi = own_v[r], wi = own_p[r] - gamma * own_v[r];
      const bool live = r < nr && g + r * G >= j;
      double acc = 0.0;
#pragma unroll
      for (int c = 0; c < kTrdCols; ++c) {
        const int l = t + kTrdThreads * c;
        if (live && l >= j && l < n) x[r][c] -= vi * wr[c] + wi * vr[c];
        acc += x[r][c] * vn[c];
        if (r == rpub && l > j && l < n) st_shared(a.Rp + (long)(j + 1) * a.ns + l, x[r][c]);
      }
      sacc[r] = (r < nr && g + r * G > j) ? acc : 0.0;
    }
    wsum8_store(sacc, rb[wv], lane);
    __syncthreads();
    double partial = 0.0;
    if (t < nr) {
      const int i = g + t * G;
      double p = 0.0;
      if (i >= j + 1) {
        double sum = 0.0;
#pragma unroll
        for (int w = 0; w < kTrdWaves; ++w) sum += rb[w][t];
        p = tj * sum;
        st_shared(a.P + (long)j * a.ns + i, p);
        partial = own_vn[t] * p;
      }
      own_p[t] = p;
      own_v[t] = own_vn[t];
    }
    if (wv == 0) {
      partial = wsum(partial);
      if (t == 0) st_shared(a.part + (long)j * a.gs + g, partial);
    }
    tau_prev = tj;
#pragma unroll
    for (int c = 0; c < kTrdCols; ++c) vr[c] = vn[c];
    if (!grid_sync(a.bar, epoch++, &flag)) return;
  }
}

int trd_grid(int n) { return (n + kTrdRows - 1) / kTrdRows; }

// Workgroup shape of the reduction by order (8 rows each): 256 threads of 2 columns up to order
// 512, 512 threads of 2 columns up to 1024, 512 threads of 4 columns up to 2048. Measured at
// 512 / 1024 / 2048: 3.8 / 7.7 / 18.4 ms for the whole solve, against 4.8 / 9.2 / 19.5 with
// 1024 threads of 2 columns at every order (idle columns, and wave sums over 16 waves), and
// 21.6 ms at 2048 with 256 threads of 8.
int trd_threads(int n) { return n <= 512 ? 256 : 512; }

const void* trd_kernel(int n) {
  if (n <= 512) return reinterpret_cast<const void*>(&sytrd_coop_kernel<256, 2>);
  if (n <= 1024) return reinterpret_cast<const void*>(&sytrd_coop_kernel<512, 2>);
  return reinterpret_cast<const void*>(&sytrd_coop_kernel<512, 4>);
}

int trd_slot(int n) { return (n + 31) / 32 * 32 + 256; }

size_t trd_ws_bytes(int n) {
  const int G = trd_grid(n), ns = trd_slot(n), gs = (G + 31) / 32 * 32 + 32;
  size_t b = 0;
  b += align_up((size_t)n * ns * sizeof(double), 256);      // P slots
  b += align_up((size_t)n * ns * sizeof(double), 256);      // Rp slots
  b += align_up((size_t)n * gs * sizeof(double), 256);      // part slots
  b += align_up(kBarWords * sizeof(unsigned), 256);         // barrier words
  b += align_up(sizeof(int), 256);                          // err
  return b;
}

// Launches the reduction; returns TNH_OK or an error code (cooperative launch refused, spin
// limit hit). Checking the spin flag needs a host read, so it is left to the caller (err).
int launch_sytrd(int n, const double* A, long lda, double* Vr, double* d, double* e, double* tau,
                 Arena& ar, int** err_out, hipStream_t st) {
  const int G = trd_grid(n), ns = trd_slot(n), gs = (G + 31) / 32 * 32 + 32;
  double* P = ar.take<double>((size_t)n * ns);
  double* Rp = ar.take<double>((size_t)n * ns);
  double* part = ar.take<double>((size_t)n * gs);
  unsigned* bar = ar.take<unsigned>(kBarWords);
  int* err = ar.take<int>(1);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "sytrd: workspace too small");
  *err_out = err;
  hipLaunchKernelGGL(bar_reset_kernel, dim3(1), dim3(256), 0, st, bar, err);
  TNH_CHECK_LAUNCH("sytrd barrier reset");
  const size_t smem = 0;
  hipError_t he;
  TrdArgs args{n, ns, gs, A, lda, Vr, d, e, tau, P, Rp, part, GridBar{bar, err}};
  void* kargs[] = {&args};
  const int nt = trd_threads(n);
  he = hipLaunchCooperativeKernel(trd_kernel(n), dim3(G), dim3(nt), kargs, (unsigned)smem, st);
  TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP,
              "sytrd: cooperative launch of %d workgroups (%zu B LDS each) refused: %s", G, smem,
              hipGetErrorString(he));
  return TNH_OK;
}

// ---- phase 2: divide and conquer on the tridiagonal ------------------------------------------
// T (order n) is padded to N = s 2^L (s <= 32) with decoupled entries far below its spectrum
// (zero couplings), so every node of level l has order s 2^l and every launch is uniform. Leaf
// b holds rows [b s, (b+1) s) with the tearing corrections |e| on its boundary diagonals and is
// solved by the batched Jacobi kernel. A node [lo, lo + K) with halves of n1 = K / 2 and coupling
// beta = e[lo + n1 - 1] is diag(T1, T2) + rho u u^T, rho = |beta|, u = (e_last; sign(beta) e_1),
// i.e. diag(Q1, Q2) (diag(D) + rho z z^T) diag(Q1, Q2)^T with z = (last row of Q1; sign row 1
// of Q2). Per level:
//   dc_prep       (workgroup per node) sort D by merging the halves, normalise z, deflate small
//                 z entries and close pairs (Givens), list the K' remaining entries
//   dc_secular    (wave per root) bisection for the roots of 1/rho + sum z_i^2 / (D_i - lam),
//                 as an offset mu from the nearer pole (Gu and Eisenstat's origin shift)
//   dc_zhat       (wave per entry) the z for which the computed roots are exact (Loewner)
//   dc_norms      (wave per eigenpair) eigenvector norms and the final ascending positions
//   dc_rows       (wave per row) M = P R (eigenvectors of the rank-one problem), row by row
//   dc_rotate     (workgroup per node) the deflation rotations on rows of M, in reverse order
//   two batched GEMMs: Qt_new = M^T-products with the halves' eigenvector blocks.
// Eigenvectors are stored as rows, node by node: the K x K block of node b (at b K K) holds
// component i of its eigenvector c at [c][i], so a level's GEMMs have uniform strides. B matrices
// are solved at once as consecutive N-blocks of one padded problem of order B N (no coupling
// between the blocks, no node spanning two of them): every launch covers all B matrices.

constexpr int kDcLeaf = 32;
constexpr int kDcTerms = kSyevMaxN / kWave;     // secular terms per lane
constexpr double kEps = 2.220446049250313e-16;

struct DcLevel {
  int K, n1, N, top;        // node order, half, padded order of one matrix, 1 on the last level
  const double* w_in;       // the halves' eigenvalues, ascending per half
  const double* Qt_in;      // the halves' eigenvector blocks (n1 x n1 each, node b's at 2b, 2b+1)
  const double* e;          // couplings (length N - 1, padded with zeros)
  double* w_out;            // N: ascending per node
  double* Mb;               // nodes x K x K
  // per-level arrays of N (node b at offset b K)
  int* perm;                // sorted position -> local index
  double* Ds;               // sorted (rotated) D
  double* zs;               // sorted normalised z, zero where deflated
  int* inv;                 // sorted position -> index among the K' kept entries, -1 deflated
  int* kept;                // kept index -> sorted position
  int* rot_j;               // rotation r of node b: rows (rot_jl[r], rot_j[r]) of sorted space
  int* rot_jl;
  double* rot_c;
  double* rot_s;
  int* cnt;                 // per node: [0] K', [1] rotations
  double* rhop;             // per node: rho' (rho |z|^2)
  int* dl;                  // deflated sorted positions, in order
  double* mu;               // per kept root j: offset from its origin
  double* dor;              // per kept root j: D' at its origin
  int* ent;                 // final column -> entity (kept roots first, then deflated)
  double* zh;               // per kept entry: zhat
  double* nrm;              // per kept root: 1 / |u_j|
};

__device__ __forceinline__ double wave_bcast_sum(double v) {
  v = wsum(v);
  return __shfl(v, 0, 64);
}

// One workgroup of 1024 threads per node.
__global__ void __launch_bounds__(1024) dc_prep_kernel(DcLevel L) {
  extern __shared__ __align__(16) char smem_raw[];
  const int K = L.K, n1 = L.n1, b = blockIdx.x, t = threadIdx.x;
  const int lo = b * K;
  double* D = reinterpret_cast<double*>(smem_raw);    // K (by local index)
  double* z = D + K;                                  // K
  double* Ds = z + K;                                 // K (sorted)
  double* zs = Ds + K;                                // K
  int* perm = reinterpret_cast<int*>(zs + K);         // K
  __shared__ double red[16];
  __shared__ double sc[2];
  const double beta = L.e[lo + n1 - 1];
  const double rho = fabs(beta), sg = beta < 0.0 ? -1.0 : 1.0;
  for (int i = t; i < K; i += blockDim.x) {
    D[i] = L.w_in[lo + i];
    z[i] = i < n1 ? L.Qt_in[(long)lo * n1 + (long)i * n1 + n1 - 1]
                  : sg * L.Qt_in[(long)lo * n1 + (long)n1 * n1 + (long)(i - n1) * n1];
    perm[i] = i;
  }
  __syncthreads();
  // merge of the two ascending halves by rank (ties: first half first); a non-finite input can
  // make ranks collide, which leaves duplicates but every index in range
  for (int i = t; i < K; i += blockDim.x) {
    const double x = D[i];
    int r;
    if (i < n1) {      // count of second-half values < x
      int a = 0, c = K - n1;
      while (a < c) {
        const int m = (a + c) >> 1;
        if (D[n1 + m] < x) a = m + 1; else c = m;
      }
      r = i + a;
    } else {           // count of first-half values <= x
      int a = 0, c = n1;
      while (a < c) {
        const int m = (a + c) >> 1;
        if (D[m] <= x) a = m + 1; else c = m;
      }
      r = (i - n1) + a;
    }
    perm[r] = i;
  }
  __syncthreads();
  double zz = 0.0, dm = 0.0;
  for (int q = t; q < K; q += blockDim.x) {
    Ds[q] = D[perm[q]];
    zs[q] = z[perm[q]];
    zz += zs[q] * zs[q];
    dm = fmax(dm, fabs(Ds[q]));
  }
  // reductions: |z|^2 and max |D|
  zz = wsum(zz);
  for (int o = 32; o >= 1; o >>= 1) dm = fmax(dm, __shfl_xor(dm, o, 64));
  if ((t & 63) == 0) {
    red[t >> 6] = zz;
  }
  __syncthreads();
  if (t == 0) {
    double s = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) s += red[w];
    sc[0] = s;
  }
  __syncthreads();
  if ((t & 63) == 0) red[t >> 6] = dm;
  __syncthreads();
  if (t == 0) {
    double m = 0.0;
    for (int w = 0; w < (int)(blockDim.x >> 6); ++w) m = fmax(m, red[w]);
    sc[1] = m;
  }
  __syncthreads();
  const double zn2 = sc[0], dmax = sc[1];
  const double rhop = rho * zn2;
  const double zinv = zn2 > 0.0 ? 1.0 / sqrt(zn2) : 0.0;
  for (int q = t; q < K; q += blockDim.x) zs[q] *= zinv;
  int* sinv = perm + K;     // K: sorted position -> kept index, negative when deflated
  int* skept = sinv + K;    // K: kept index -> sorted position
  __shared__ int snk, snr;
  __syncthreads();
  if (t == 0) {
    // deflation, sequential in sorted order (tol as in LAPACK's dlaed2); the last kept entry's
    // z and D ride in registers
    const double tol = 8.0 * kEps * fmax(dmax, rhop);
    int nk = 0, nr = 0, jl = -1;
    double zl = 0.0, dl = 0.0;
    // entry j + 1 is read while entry j is processed (an iteration writes entries j and jl < j
    // only), so the LDS latency of the next read overlaps this iteration's work
    double zn = zs[0], dn = Ds[0];
    for (int j = 0; j < K; ++j) {
      double zj = zn, dj = dn;
      if (j + 1 < K) {
        zn = zs[j + 1];
        dn = Ds[j + 1];
      }
      if (rhop * fabs(zj) <= tol) {
        zs[j] = 0.0;
        sinv[j] = -1;
        continue;
      }
      // close pair: |(d_j - d_l) c s| <= tol with c = z_j / t, s = -z_l / t, t = |(z_j, z_l)|
      if (jl >= 0 && fabs((dj - dl) * zj * zl) <= tol * (zj * zj + zl * zl)) {
        const double tt = sqrt(zj * zj + zl * zl), it = 1.0 / tt;
        const double c = zj * it, sn = -zl * it;
        const double a2 = dl * c * c + dj * sn * sn, b2 = dl * sn * sn + dj * c * c;
        Ds[jl] = a2;
        Ds[j] = b2;
        dj = b2;
        zj = tt;
        zs[j] = tt;
        zs[jl] = 0.0;
        L.rot_jl[lo + nr] = jl;
        L.rot_j[lo + nr] = j;
        L.rot_c[lo + nr] = c;
        L.rot_s[lo + nr] = sn;
        ++nr;
        sinv[jl] = -1;     // jl leaves the kept list (it was appended last)
        --nk;
      }
      sinv[j] = nk;
      skept[nk++] = j;
      jl = j;
      zl = zj;
      dl = dj;
    }
    int nd = 0;
    for (int j = 0; j < K; ++j)
      if (sinv[j] < 0) {
        sinv[j] = -1 - nd;
        skept[nk + nd++] = j;      // deflated positions follow the kept ones
      }
    snk = nk;
    snr = nr;
  }
  __syncthreads();
  const int nk = snk;
  for (int q = t; q < K; q += blockDim.x) {
    L.perm[lo + q] = perm[q];
    L.Ds[lo + q] = Ds[q];
    L.zs[lo + q] = zs[q];
    L.inv[lo + q] = sinv[q];
    if (q < nk) L.kept[lo + q] = skept[q];
    else L.dl[lo + q - nk] = skept[q];
  }
  if (t == 0) {
    L.cnt[2 * b] = nk;
    L.cnt[2 * b + 1] = snr;
    L.rhop[b] = rhop;
  }
}

// One wave per kept root j of node blockIdx.y (4 waves per block). The root of
// f(mu) = 1/rho + sum z_i^2 / (delta_i - mu) (delta_i = D_i - D_origin) is bracketed and refined
// by the rational model of Bunch, Nielsen and Sorensen: the poles left and right of the interval
// are each replaced by c + s / (delta - x), matched in value and slope at the current iterate,
// and the model's root in the bracket is the next iterate (quadratic convergence); a step that
// leaves the bracket is replaced by bisection, and every evaluation tightens the bracket.
__global__ void __launch_bounds__(256) dc_secular_kernel(DcLevel L) {
  const int K = L.K, b = blockIdx.y, lane = threadIdx.x & 63;
  const int j = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lo = b * K, nk = L.cnt[2 * b];
  if (j >= nk) return;
  const int* kept = L.kept + lo;
  const double invr = 1.0 / L.rhop[b];
  double dv[kDcTerms], z2[kDcTerms];
#pragma unroll
  for (int c = 0; c < kDcTerms; ++c) {
    const int i = lane + kWave * c;
    dv[c] = 0.0;
    z2[c] = 0.0;
    if (i < nk) {
      const int q = kept[i];
      dv[c] = L.Ds[lo + q];
      z2[c] = L.zs[lo + q] * L.zs[lo + q];
    }
  }
  const double dj = L.Ds[lo + kept[j]];
  const bool last = j == nk - 1;
  const double gap = last ? 0.0 : L.Ds[lo + kept[j + 1]] - dj;
  int o = j;
  double a, c2;   // bracket of mu
  if (!last) {
    double f = 0.0;
#pragma unroll
    for (int c = 0; c < kDcTerms; ++c) f += z2[c] / ((dv[c] - dj) - 0.5 * gap);
    f = invr + wave_bcast_sum(f);
    if (f >= 0.0) {
      a = 0.0;
      c2 = 0.5 * gap;
    } else {
      o = j + 1;
      a = -0.5 * gap;
      c2 = 0.0;
    }
  } else {
    double zz = 0.0;
#pragma unroll
    for (int c = 0; c < kDcTerms; ++c) zz += z2[c];
    a = 0.0;
    c2 = wave_bcast_sum(zz) / invr;
  }
  const double dorg = o == j ? dj : L.Ds[lo + kept[o]];
#pragma unroll
  for (int c = 0; c < kDcTerms; ++c) dv[c] -= dorg;    // pole offsets from the origin
  const double dl = dj - dorg, dr = last ? 0.0 : dl + gap;   // the interval's poles
  double mu = 0.5 * (a + c2);
  for (int it = 0; it < 200; ++it) {
    // psi (poles <= j), phi (poles > j), and their slopes at mu
    double ps = 0.0, pd = 0.0, fs = 0.0, fd = 0.0;
#pragma unroll
    for (int c = 0; c < kDcTerms; ++c) {
      const int i = lane + kWave * c;
      const double inv = 1.0 / (dv[c] - mu);
      const double tt = z2[c] * inv, td = tt * inv;
      if (i <= j) {
        ps += tt;
        pd += td;
      } else {
        fs += tt;
        fd += td;
      }
    }
    ps = wave_bcast_sum(ps);
    pd = wave_bcast_sum(pd);
    fs = wave_bcast_sum(fs);
    fd = wave_bcast_sum(fd);
    const double f = invr + ps + fs;
    if (f > 0.0) c2 = mu; else a = mu;
    // f is known to within eps * err (dlaed4's bound); below that the root is found
    const double err = 8.0 * (fs - ps) + 2.0 * invr + 3.0 * fabs(mu) * (pd + fd);
    if (fabs(f) <= kEps * err || c2 - a <= 2.0 * kEps * fmax(fabs(a), fabs(c2))) break;
    // model: cc + q / (dl - x) + s / (dr - x) = 0, x = mu + eta
    const double Dl = dl - mu, Dr = dr - mu;
    const double q = pd * Dl * Dl;
    double nx;
    if (last) {
      const double cc = invr + ps - pd * Dl;     // cc + q / (Dl - eta) = 0
      nx = cc > 0.0 ? mu + Dl + q / cc : 0.5 * (a + c2);
    } else {
      const double sr = fd * Dr * Dr;
      const double cc = invr + (ps - pd * Dl) + (fs - fd * Dr);
      // cc eta^2 - B eta + C = 0, C = Dl Dr f
      const double B = cc * (Dl + Dr) + q + sr, C = Dl * Dr * f;
      const double disc = B * B - 4.0 * cc * C;
      double eta;
      if (disc < 0.0) eta = 0.5 * (a + c2) - mu;
      else if (cc == 0.0) eta = B != 0.0 ? C / B : 0.5 * (a + c2) - mu;
      else eta = 2.0 * C / (B + copysign(sqrt(disc), B));
      nx = mu + eta;
    }
    if (!(nx > a && nx < c2)) nx = 0.5 * (a + c2);
    if (nx == mu) break;
    mu = nx;
  }
  if (lane == 0) {
    L.mu[lo + j] = mu;
    L.dor[lo + j] = dorg;
  }
}

// zhat_i (Loewner): zhat_i^2 = -(D_i - lam_i) prod_{j != i} (D_i - lam_j) / (D_i - D_j) / rho',
// with D_i - lam_j = (D_i - D_origin(j)) - mu_j; the product is kept as mantissa and exponent.
__global__ void __launch_bounds__(256) dc_zhat_kernel(DcLevel L) {
  const int K = L.K, b = blockIdx.y, lane = threadIdx.x & 63;
  const int i = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lo = b * K, nk = L.cnt[2 * b];
  if (i >= nk) return;
  const int* kept = L.kept + lo;
  const double di = L.Ds[lo + kept[i]];
  double m = 1.0;
  int ex = 0;
  for (int j = lane; j < nk; j += kWave) {
    const double q = (di - L.dor[lo + j]) - L.mu[lo + j];
    const double f = j == i ? q : q / (di - L.Ds[lo + kept[j]]);
    int e2;
    m = frexp(m * f, &e2);
    ex += e2;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) {
    const double m2 = __shfl_xor(m, o, 64);
    const int e2 = __shfl_xor(ex, o, 64);
    int e3;
    m = frexp(m * m2, &e3);
    ex += e2 + e3;
  }
  m = __shfl(m, 0, 64);
  ex = __shfl(ex, 0, 64);
  if (lane == 0) {
    double v = -m / L.rhop[b];          // zhat^2 = v 2^ex, v in (0, 1] when the roots interlace
    if (ex & 1) {
      v *= 2.0;
      ex -= 1;
    }
    const double zh = v > 0.0 ? ldexp(sqrt(v), ex / 2) : 0.0;
    L.zh[lo + i] = L.zs[lo + kept[i]] < 0.0 ? -zh : zh;
  }
}

// Entity e of node b: a kept root (e < K') or the (e - K')-th deflated entry. Its eigenvector
// norm (roots), its value and its rank among all K values (ascending; ties by entity) give the
// final column: rank, or K - 1 - rank on the last level (descending order).
__global__ void __launch_bounds__(256) dc_norms_kernel(DcLevel L) {
  const int K = L.K, b = blockIdx.y, lane = threadIdx.x & 63;
  const int e = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lo = b * K, nk = L.cnt[2 * b];
  if (e >= K) return;
  const int* kept = L.kept + lo;
  double val;
  if (e < nk) {
    const double dor = L.dor[lo + e], mu = L.mu[lo + e];
    val = dor + mu;
    double s = 0.0;
    for (int i = lane; i < nk; i += kWave) {
      const double u = L.zh[lo + i] / ((L.Ds[lo + kept[i]] - dor) - mu);
      s += u * u;
    }
    s = wave_bcast_sum(s);
    if (lane == 0) L.nrm[lo + e] = s > 0.0 ? 1.0 / sqrt(s) : 0.0;
  } else {
    val = L.Ds[lo + L.dl[lo + e - nk]];
  }
  int r = 0;
  for (int f = lane; f < K; f += kWave) {
    const double vf = f < nk ? L.dor[lo + f] + L.mu[lo + f] : L.Ds[lo + L.dl[lo + f - nk]];
    r += (vf < val || (vf == val && f < e)) ? 1 : 0;
  }
#pragma unroll
  for (int o = 32; o >= 1; o >>= 1) r += __shfl_xor(r, o, 64);
  if (lane == 0) {
    L.w_out[lo + r] = val;
    L.ent[lo + (L.top ? K - 1 - r : r)] = e;
  }
}

// Row perm[q] of M for sorted position q: the eigenvector entries of kept row inv[q] (or the unit
// row of a deflated entry) at every final column < ncol.
__global__ void __launch_bounds__(256) dc_rows_kernel(DcLevel L, int ncol) {
  const int K = L.K, b = blockIdx.y, lane = threadIdx.x & 63;
  const int q = blockIdx.x * 4 + (threadIdx.x >> 6);
  const int lo = b * K, nk = L.cnt[2 * b];
  if (q >= K) return;
  const int iv = L.inv[lo + q];
  double* row = L.Mb + (long)b * K * ncol + (long)L.perm[lo + q] * ncol;
  if (iv < 0) {
    const int me = nk + (-1 - iv);
    for (int c = lane; c < ncol; c += kWave) row[c] = L.ent[lo + c] == me ? 1.0 : 0.0;
    return;
  }
  const double zi = L.zh[lo + iv], di = L.Ds[lo + q];
  for (int c = lane; c < ncol; c += kWave) {
    const int e = L.ent[lo + c];   // in [0, K) unless a non-finite value broke the ranking
    row[c] = (e >= 0 && e < nk) ? zi / ((di - L.dor[lo + e]) - L.mu[lo + e]) * L.nrm[lo + e] : 0.0;
  }
}

// The deflation rotations of node blockIdx.y on rows of M, last first: one thread per column
// applies all of them in order, keeping the row that the next rotation of a chain takes in a
// register, so a chain costs one load and one store per rotation and no barrier.
__global__ void __launch_bounds__(256) dc_rotate_kernel(DcLevel L, int ncol) {
  const int K = L.K, b = blockIdx.y, lo = b * K;
  const int col = blockIdx.x * blockDim.x + threadIdx.x;
  const int nr = L.cnt[2 * b + 1];
  if (col >= ncol || nr == 0) return;
  double* Mc = L.Mb + (long)b * K * ncol + col;
  int held = -1;
  double hv = 0.0;
  for (int r = nr - 1; r >= 0; --r) {
    const int ra = L.perm[lo + L.rot_jl[lo + r]], rb = L.perm[lo + L.rot_j[lo + r]];
    const double c = L.rot_c[lo + r], sn = L.rot_s[lo + r];
    const double x = ra == held ? hv : Mc[(long)ra * ncol];
    const double y = rb == held ? hv : Mc[(long)rb * ncol];
    if (held >= 0 && held != ra && held != rb) Mc[(long)held * ncol] = hv;
    Mc[(long)rb * ncol] = sn * x + c * y;
    held = ra;
    hv = c * x - sn * y;
  }
  Mc[(long)held * ncol] = hv;
}

// Padding and leaves. One workgroup per matrix (blockIdx.x; d, e at strides sd, se): the bound
// |T| <= max|d| + 2 max|e|, the padded d and e (decoupled entries at -3 bound, no coupling to the
// next matrix) at dp, ep + b N, then leaf matrices with the tearing corrections.
__global__ void __launch_bounds__(1024) dc_pad_kernel(int n, int N, const double* d, long sd,
                                                      const double* e, long se, double* dp,
                                                      double* ep) {
  __shared__ double red[16];
  d += blockIdx.x * sd;
  e += blockIdx.x * se;
  dp += (long)blockIdx.x * N;
  ep += (long)blockIdx.x * N;
  double m = 0.0;
  for (int i = threadIdx.x; i < n; i += blockDim.x) {
    m = fmax(m, fabs(d[i]));
    if (i < n - 1) m = fmax(m, 2.0 * fabs(e[i]));
  }
  for (int o = 32; o >= 1; o >>= 1) m = fmax(m, __shfl_xor(m, o, 64));
  if ((threadIdx.x & 63) == 0) red[threadIdx.x >> 6] = m;
  __syncthreads();
  double bound = 0.0;
  for (int w = 0; w < (int)(blockDim.x >> 6); ++w) bound = fmax(bound, red[w]);
  const double dummy = bound > 0.0 ? -3.0 * bound : -1.0;
  for (int i = threadIdx.x; i < N; i += blockDim.x) {
    dp[i] = i < n ? d[i] : dummy;
    ep[i] = i < n - 1 ? e[i] : 0.0;
  }
}

__global__ void dc_leaves_kernel(int s, int N, const double* dp, const double* ep, double* A) {
  const int b = blockIdx.x, lo = b * s;
  double* Ab = A + (long)b * s * s;
  for (int idx = threadIdx.x; idx < s * s; idx += blockDim.x) {
    const int i = idx / s, j = idx - i * s;
    double v = 0.0;
    if (i == j) {
      v = dp[lo + i];
      if (i == 0 && lo > 0) v -= fabs(ep[lo - 1]);
      if (i == s - 1 && lo + s < N) v -= fabs(ep[lo + s - 1]);
    } else if (j == i + 1) {
      v = ep[lo + i];
    } else if (i == j + 1) {
      v = ep[lo + j];
    }
    Ab[idx] = v;
  }
}

// Leaf block b of Qt: row c = eigenvector c of leaf b (Qt[b s s + c s + i] = Z_b[i][c]).
__global__ void dc_scatter_leaves_kernel(int s, const double* Z, double* Qt) {
  const int b = blockIdx.x;
  const double* Zb = Z + (long)b * s * s;
  double* Qb = Qt + (long)b * s * s;
  for (int idx = threadIdx.x; idx < s * s; idx += blockDim.x) {
    const int c = idx / s, i = idx - c * s;
    Qb[idx] = Zb[i * s + c];
  }
}


// ---- phase 3: back-transformation ------------------------------------------------------------
// Z <- Q Z with Q = Q_0 Q_1 ... (blocks of kBtNb reflectors, Q_b = I - V T V^T, T upper
// triangular from the recurrence T[0:c, c] = -tau_c T[0:c, 0:c] (V[:, 0:c]^T v_c), T[c][c] =
// tau_c), applied last block first: Y = V^T Z, Y' = -T Y, Z += V Y' (three K-major GEMMs).
constexpr int kBtNb = 64;

// Vc[i][j] = Vr[j][i] for reflectors j < n - 1, zero for n - 1 <= j < NV (row stride NV).
__global__ void transpose_kernel(int n, int NV, const double* __restrict__ src,
                                 double* __restrict__ dst) {
  __shared__ double tile[32][33];
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;   // bx: column block of src (i), by: j
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int j = by + r, i = bx + threadIdx.x;
    tile[r][threadIdx.x] = (j < n - 1 && i < n) ? src[(long)j * n + i] : 0.0;
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int i = bx + r, j = by + threadIdx.x;
    if (i < n && j < NV) dst[(long)i * NV + j] = tile[threadIdx.x][r];
  }
}

// Panel b (one workgroup): from S = V_b^T V_b (nb x nb) and tau, writes A_b = -T^T (row p, col q
// = -T[q][p]), the left operand of Y' = -T Y in the K-major GEMM; S, tau and negTt of matrix
// blockIdx.y at y sS, y stau, y sS.
__global__ void __launch_bounds__(256) larft_kernel(int nrefl, const double* S, const double* tau,
                                                    double* negTt, long sS, long stau) {
  __shared__ double T[kBtNb][kBtNb + 1];
  __shared__ double tmp[kBtNb];
  const int b = blockIdx.x, jb = b * kBtNb, nb = min(kBtNb, nrefl - jb), t = threadIdx.x;
  S += blockIdx.y * sS;       // matrix blockIdx.y
  tau += blockIdx.y * stau;
  negTt += blockIdx.y * sS;
  const double* Sb = S + (long)b * kBtNb * kBtNb;
  for (int idx = t; idx < kBtNb * kBtNb; idx += blockDim.x) T[idx / kBtNb][idx % kBtNb] = 0.0;
  __syncthreads();
  for (int c = 0; c < nb; ++c) {
    const double tc = tau[jb + c];
    // tmp[p] = sum_{q < c} T[p][q] S[q][c] for p < c (T upper: q >= p)
    if (t < c) {
      double s = 0.0;
      for (int q = t; q < c; ++q) s += T[t][q] * Sb[q * kBtNb + c];
      tmp[t] = s;
    }
    __syncthreads();
    if (t < c) T[t][c] = -tc * tmp[t];
    if (t == 0) T[c][c] = tc;
    __syncthreads();
  }
  double* out = negTt + (long)b * kBtNb * kBtNb;
  for (int idx = t; idx < kBtNb * kBtNb; idx += blockDim.x) {
    const int p = idx / kBtNb, q = idx % kBtNb;
    out[idx] = -T[q][p];
  }
}

__global__ void syev_status_kernel(const int* err, const int* sweeps, int nleaf, int* info) {
  int bad = 0;
  for (int i = threadIdx.x; i < nleaf; i += blockDim.x) bad |= sweeps[i] < 0;
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) *info = (*err ? 1 : 0) | (bad ? 2 : 0);
}

struct SyevPlan {
  int n, k, L, s, N, npan, B;   // B: matrices solved together by the divide and conquer
};

SyevPlan syev_plan(int n, int k, int B = 1) {
  SyevPlan p;
  p.n = n;
  p.k = k;
  p.B = B;
  p.L = 1;
  while ((n + (1 << p.L) - 1) >> p.L > kDcLeaf) ++p.L;
  p.s = (n + (1 << p.L) - 1) >> p.L;
  p.N = p.s << p.L;
  p.npan = (n - 1 + kBtNb - 1) / kBtNb;
  return p;
}

// Workspace pieces, in carving order (one function, used for both the size and the carve).
template <typename F>
void syev_pieces(const SyevPlan& p, F&& f) {
  const size_t n = p.n, N = p.N, k = p.k, npan = p.npan, NB = (size_t)p.B * p.N;
  f("trd", tnh::trd_ws_bytes(p.n));
  f("Vr", n * n * 8);
  f("Vc", n * npan * kBtNb * 8);
  f("d", n * 8);
  f("e", n * 8);
  f("tau", n * 8);
  f("dp", NB * 8);
  f("ep", NB * 8);
  f("QtA", NB * N * 8);
  f("QtB", NB * N * 8);
  f("Mb", NB * N * 8);
  f("leafA", NB * p.s * 8);
  f("leafZ", NB * p.s * 8);
  f("wA", NB * 8);
  f("wB", NB * 8);
  f("sweeps", (NB / p.s) * 4);
  for (const char* nm : {"perm", "inv", "kept", "dl", "rot_j", "rot_jl", "ent"}) f(nm, NB * 4);
  for (const char* nm : {"Ds", "zs", "rot_c", "rot_s", "mu", "dor", "zh", "nrm"}) f(nm, NB * 8);
  f("cnt", NB * 8);
  f("rhop", NB * 8);
  f("Zw", NB * k * 8);
  f("S", npan * kBtNb * kBtNb * 8);
  f("negTt", npan * kBtNb * kBtNb * 8);
  f("Y", kBtNb * k * 8);
  f("Y2", kBtNb * k * 8);
  f("info", 256);
}

size_t syev_ws_bytes(const SyevPlan& p) {
  size_t tot = 0;
  syev_pieces(p, [&](const char*, size_t b) { tot += align_up(b, 256); });
  return tot + 256;
}

int gemm(int M, int N, int K, const double* A, long lda, const double* B, long ldb, double* C,
         long ldc, int batch, long sA, long sB, long sC, int acc, hipStream_t st) {
  return tnh_gemm_tn(TNH_F64, M, N, K, A, lda, B, ldb, C, ldc, batch, sA, sB, sC, acc, 0, st);
}

// The workspace of one solve, carved by syev_pieces.
struct SyevBufs {
  char* trd;
  double *Vr, *Vc, *d, *e, *tau, *dp, *ep, *Qt[2], *Mb, *leafA, *leafZ, *wv[2];
  int* sweeps;
  DcLevel Lv;
  double *Zw, *S, *negTt, *Y, *Y2;
};

SyevBufs syev_carve(const SyevPlan& p, void* ws) {
  std::vector<char*> ptrs;
  char* base = reinterpret_cast<char*>(align_up(reinterpret_cast<size_t>(ws), 256));
  size_t off = 0;
  syev_pieces(p, [&](const char*, size_t b) {
    ptrs.push_back(base + off);
    off += align_up(b, 256);
  });
  int q = 0;
  auto next = [&]() { return ptrs[q++]; };
  auto dbl = [&]() { return reinterpret_cast<double*>(next()); };
  auto itg = [&]() { return reinterpret_cast<int*>(next()); };
  SyevBufs b;
  b.trd = next();
  b.Vr = dbl();
  b.Vc = dbl();
  b.d = dbl();
  b.e = dbl();
  b.tau = dbl();
  b.dp = dbl();
  b.ep = dbl();
  b.Qt[0] = dbl();
  b.Qt[1] = dbl();
  b.Mb = dbl();
  b.leafA = dbl();
  b.leafZ = dbl();
  b.wv[0] = dbl();
  b.wv[1] = dbl();
  b.sweeps = itg();
  b.Lv.perm = itg();
  b.Lv.inv = itg();
  b.Lv.kept = itg();
  b.Lv.dl = itg();
  b.Lv.rot_j = itg();
  b.Lv.rot_jl = itg();
  b.Lv.ent = itg();
  b.Lv.Ds = dbl();
  b.Lv.zs = dbl();
  b.Lv.rot_c = dbl();
  b.Lv.rot_s = dbl();
  b.Lv.mu = dbl();
  b.Lv.dor = dbl();
  b.Lv.zh = dbl();
  b.Lv.nrm = dbl();
  b.Lv.cnt = itg();
  b.Lv.rhop = dbl();
  b.Zw = dbl();
  b.S = dbl();
  b.negTt = dbl();
  b.Y = dbl();
  b.Y2 = dbl();
  return b;
}

// Divide and conquer on the tridiagonals (d + m sd, e + m se) of order p.n of p.B matrices: all
// eigenvalues ascending into w + m sw, the eigenvectors of the k largest into rows 0..n-1 of
// b.Zw + m N k (n x k, column c the c-th largest).
int run_dc(const SyevPlan& p, SyevBufs& b, const double* d, long sd, const double* e, long se,
           double* w, long sw, hipStream_t st) {
  const int n = p.n, N = p.N, k = p.k, s = p.s, B = p.B, NB = B * N;
  double* const* Qt = b.Qt;
  double* const* wv = b.wv;
  DcLevel& Lv = b.Lv;
  double* Mb = b.Mb;
  double* Zw = b.Zw;
  double *dp = b.dp, *ep = b.ep, *leafA = b.leafA, *leafZ = b.leafZ;
  int* sweeps = b.sweeps;
  int rc;
  hipLaunchKernelGGL(dc_pad_kernel, dim3(B), dim3(1024), 0, st, n, N, d, sd, e, se, dp, ep);
  hipLaunchKernelGGL(dc_leaves_kernel, dim3(NB / s), dim3(256), 0, st, s, NB, dp, ep, leafA);
  TNH_CHECK_LAUNCH("syev leaves");
  rc = tnh_eigh_jacobi(TNH_F64, NB / s, s, leafA, s, (long)s * s, nullptr, 0, wv[0], s, leafZ, s,
                       (long)s * s, 60, 2e-16, sweeps, st);
  if (rc != TNH_OK) return rc;
  hipLaunchKernelGGL(dc_scatter_leaves_kernel, dim3(NB / s), dim3(256), 0, st, s, leafZ, Qt[0]);
  TNH_CHECK_LAUNCH("syev scatter");
  int cur = 0;
  Lv.N = N;
  Lv.e = ep;
  Lv.Mb = Mb;
  for (int l = 1; l <= p.L; ++l) {
    const int K = s << l, nodes = NB / K, top = l == p.L;
    Lv.K = K;
    Lv.n1 = K / 2;
    Lv.top = top;
    Lv.w_in = wv[cur];
    Lv.w_out = wv[cur ^ 1];
    Lv.Qt_in = Qt[cur];
    const size_t psm = (size_t)K * (4 * sizeof(double) + 3 * sizeof(int));
    hipError_t he = hipFuncSetAttribute(reinterpret_cast<const void*>(&dc_prep_kernel),
                                        hipFuncAttributeMaxDynamicSharedMemorySize, (int)psm);
    TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP, "syev: prep LDS %zu: %s", psm,
                hipGetErrorString(he));
    hipLaunchKernelGGL(dc_prep_kernel, dim3(nodes), dim3(1024), psm, st, Lv);
    const dim3 wg((K + 3) / 4, nodes);
    hipLaunchKernelGGL(dc_secular_kernel, wg, dim3(256), 0, st, Lv);
    hipLaunchKernelGGL(dc_zhat_kernel, wg, dim3(256), 0, st, Lv);
    hipLaunchKernelGGL(dc_norms_kernel, wg, dim3(256), 0, st, Lv);
    const int ncol = top ? k : K;
    hipLaunchKernelGGL(dc_rows_kernel, wg, dim3(256), 0, st, Lv, ncol);
    hipLaunchKernelGGL(dc_rotate_kernel, dim3((ncol + 255) / 256, nodes), dim3(256), 0, st, Lv,
                       ncol);
    TNH_CHECK_LAUNCH("syev merge");
    const int n1 = K / 2;
    const long sK = (long)K * K, sH = sK / 2;   // a node's block, its two halves' blocks
    if (!top) {
      double* Qo = Qt[cur ^ 1];
      rc = gemm(K, n1, n1, Mb, K, Qt[cur], n1, Qo, K, nodes, sK, sH, sK, 0, st);
      if (rc != TNH_OK) return rc;
      rc = gemm(K, n1, n1, Mb + (long)n1 * K, K, Qt[cur] + (long)n1 * n1, n1, Qo + n1, K, nodes,
                sK, sH, sK, 0, st);
      if (rc != TNH_OK) return rc;
    } else {
      const long sM = (long)K * k, sZ = (long)N * k;
      rc = gemm(n1, k, n1, Qt[cur], n1, Mb, k, Zw, k, nodes, sH, sM, sZ, 0, st);
      if (rc != TNH_OK) return rc;
      rc = gemm(n1, k, n1, Qt[cur] + (long)n1 * n1, n1, Mb + (long)n1 * k, k, Zw + (long)n1 * k,
                k, nodes, sH, sM, sZ, 0, st);
      if (rc != TNH_OK) return rc;
    }
    cur ^= 1;
  }
  // eigenvalues: each matrix's top-level ascending values above its N - n padding entries
  hipError_t he = hipMemcpy2DAsync(w, (size_t)sw * sizeof(double), wv[cur] + (N - n),
                                   (size_t)N * sizeof(double), (size_t)n * sizeof(double),
                                   (size_t)B, hipMemcpyDeviceToDevice, st);
  TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP, "syev: eigenvalue copy: %s", hipGetErrorString(he));
  return TNH_OK;
}

int syev_top(const SyevPlan& p, const double* A, long lda, double* w, double* Z, long ldz,
             void* ws, size_t ws_bytes, int* info, hipStream_t st) {
  (void)ws_bytes;
  SyevBufs b = syev_carve(p, ws);
  const int n = p.n, N = p.N, k = p.k, s = p.s;
  double *Vr = b.Vr, *Vc = b.Vc, *tau = b.tau, *Zw = b.Zw, *S = b.S, *negTt = b.negTt;
  double *Y = b.Y, *Y2 = b.Y2;

  // 1. tridiagonalisation
  Arena tar(b.trd, trd_ws_bytes(n));
  int* err = nullptr;
  int rc = launch_sytrd(n, A, lda, Vr, b.d, b.e, tau, tar, &err, st);
  if (rc != TNH_OK) return rc;

  // 2. divide and conquer
  rc = run_dc(p, b, b.d, n, b.e, n, w, n, st);
  if (rc != TNH_OK) return rc;

  // 3. back-transformation of the k vectors (rows 0..n-1 of Zw)
  const int NV = p.npan * kBtNb;
  hipLaunchKernelGGL(transpose_kernel, dim3((n + 31) / 32, (NV + 31) / 32), dim3(32, 8), 0, st, n,
                     NV, Vr, Vc);
  TNH_CHECK_LAUNCH("syev transpose");
  const int nrefl = n - 1;
  rc = gemm(kBtNb, kBtNb, n, Vc, NV, Vc, NV, S, kBtNb, p.npan, kBtNb, kBtNb,
            (long)kBtNb * kBtNb, 0, st);
  if (rc != TNH_OK) return rc;
  hipLaunchKernelGGL(larft_kernel, dim3(p.npan), dim3(256), 0, st, nrefl, S, tau, negTt, 0L, 0L);
  TNH_CHECK_LAUNCH("syev larft");
  for (int bb = p.npan - 1; bb >= 0; --bb) {
    const int jb = bb * kBtNb, nb = std::min(kBtNb, nrefl - jb), rows = n - jb;
    rc = gemm(nb, k, rows, Vc + (long)jb * NV + jb, NV, Zw + (long)jb * k, k, Y, k, 1, 0, 0, 0, 0,
              st);
    if (rc != TNH_OK) return rc;
    rc = gemm(nb, k, nb, negTt + (long)bb * kBtNb * kBtNb, kBtNb, Y, k, Y2, k, 1, 0, 0, 0, 0,
              st);
    if (rc != TNH_OK) return rc;
    rc = gemm(rows, k, nb, Vr + (long)jb * n + jb, n, Y2, k, Zw + (long)jb * k, k, 1, 0, 0, 0, 1,
              st);
    if (rc != TNH_OK) return rc;
  }
  hipError_t he = hipMemcpy2DAsync(Z, (size_t)ldz * sizeof(double), Zw, (size_t)k * sizeof(double),
                                   (size_t)k * sizeof(double), (size_t)n, hipMemcpyDeviceToDevice,
                                   st);
  TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP, "syev: vector copy: %s", hipGetErrorString(he));
  if (info) {
    hipLaunchKernelGGL(syev_status_kernel, dim3(1), dim3(256), 0, st, err, b.sweeps, N / s, info);
    TNH_CHECK_LAUNCH("syev status");
  }
  return TNH_OK;
}

// ---- complex Hermitian matrices ----------------------------------------------------------------
// A = Q T Q^H with T real symmetric tridiagonal (LAPACK's zhetd2 in lower form: H_j = I - tau_j
// v_j v_j^H, H_j^H (alpha; x) = (beta; 0) with beta real, so every subdiagonal entry of T is
// real), then the real divide and conquer on T, then Z = Q Z_T by complex WY blocks. The
// reduction is the real kernel's scheme (rows in registers, one grid barrier per column) in
// complex arithmetic, with workgroups of 256 or 512 threads holding 8 rows x 2 columns per thread
// (order <= 1024), and several matrices per launch: workgroup b serves matrix b / G as its workgroup
// b % G, each matrix with its own barrier words and exchange slots.
constexpr int kHCols = 2;
constexpr int kHRows = 8;
constexpr int kHeevMaxN = 512 * kHCols;   // 1024 (512 threads; 256 up to order 512)

__device__ __forceinline__ c128 cmul(c128 a, c128 b) { return a * b; }
__device__ __forceinline__ c128 cscale(double s, c128 a) { return {s * a.re, s * a.im}; }
__device__ __forceinline__ double cabs2(c128 a) { return a.re * a.re + a.im * a.im; }
__device__ __forceinline__ void st_shared_c(c128* p, c128 v) {
  double* q = reinterpret_cast<double*>(p);
  st_shared(q, v.re);
  st_shared(q + 1, v.im);
}

struct HtrdArgs {
  int n, ns, gs, G;          // order, slot lengths, workgroups per matrix
  const c128* A;             // matrix m at A + m sA, leading dimension lda
  long lda, sA;
  c128* Vr;                  // n x n per matrix (stride n n): row j = v_j
  double* d;                 // n per matrix
  double* e;                 // n per matrix
  c128* tau;                 // n per matrix
  c128* P;                   // per matrix n x ns slots: tau_j A v_j
  c128* Rp;                  // per matrix n x ns slots: row j before its update
  c128* part;                // per matrix n x gs: per-workgroup partials p^H v
  unsigned* bar;             // per matrix kBarWords
  int* err;
};

// 256-thread workgroups are held to three per CU (168 VGPRs, a few spilled): twelve matrices of
// order 512 per launch instead of eight, 10 % less time for 32 of them.
template <int kHThreads>
__global__ void __launch_bounds__(kHThreads, kHThreads == 256 ? 3 : 1)
    hetrd_coop_kernel(HtrdArgs a) {
  constexpr int kHWaves = kHThreads / 64;
  __shared__ double red[2 * kHWaves];
  __shared__ double red8[2][kHWaves][2 * kHRows];
  __shared__ double ov_re[kHRows], ov_im[kHRows], op_re[kHRows], op_im[kHRows];
  __shared__ double on_re[kHRows], on_im[kHRows];
  __shared__ double vnx_re, vnx_im;   // v_{j-1}[j+1]
  __shared__ int flag;
  const int n = a.n, G = a.G, m = blockIdx.x / G, g = blockIdx.x % G, t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int nr = (n - g + G - 1) / G;
  const long nn = (long)n * n, sl = (long)n * a.ns, sp = (long)n * a.gs;
  const c128* A = a.A + m * a.sA;
  c128* Vr = a.Vr + m * nn;
  double* dd = a.d + (long)m * n;
  double* ee = a.e + (long)m * n;
  c128* tt = a.tau + (long)m * n;
  c128* P = a.P + m * sl;
  c128* Rp = a.Rp + m * sl;
  c128* part = a.part + m * sp;
  const GridBar bar{a.bar + (size_t)m * kBarWords, a.err};
  int par = 0, par8 = 0;

  c128 x[kHRows][kHCols];
#pragma unroll
  for (int r = 0; r < kHRows; ++r)
#pragma unroll
    for (int c = 0; c < kHCols; ++c) {
      const int l = t + kHThreads * c;
      x[r][c] = (r < nr && l < n) ? A[(long)(g + r * G) * a.lda + l] : c128{};
    }
  if (n == 1) {
    if (g == 0 && t == 0) dd[0] = x[0][0].re;
    return;
  }
  if (g == 0)
#pragma unroll
    for (int c = 0; c < kHCols; ++c) {
      const int l = t + kHThreads * c;
      if (l < n) st_shared_c(Rp + l, x[0][c]);
    }
  if (t < kHRows) {
    ov_re[t] = ov_im[t] = op_re[t] = op_im[t] = 0.0;
  }
  unsigned epoch = 1;
  if (!grid_sync_of(bar, epoch++, &flag, G, g)) return;

  c128 vr[kHCols];   // v_{j-1} at this thread's columns
  c128 tau_prev{};
  for (int j = 0; j < n - 1; ++j) {
    const c128* Rj = Rp + (long)j * a.ns;
    c128 R[kHCols], wr[kHCols];
    c128 aw{}, wj{};
    double dj;
    c128 alpha;
    if (j > 0) {
      const c128* Pj = P + (long)(j - 1) * a.ns;
      const c128* pj = part + (long)(j - 1) * a.gs;
      c128 ps{};
#pragma unroll
      for (int q = 0; q < kBarMaxG / 64; ++q) {
        const int kk = lane + 64 * q;
        if (kk < G) ps = ps + pj[kk];
      }
      c128 Pl[kHCols];
#pragma unroll
      for (int c = 0; c < kHCols; ++c) {
        const int l = t + kHThreads * c;
        Pl[c] = l >= j && l < n ? Pj[l] : c128{};
        R[c] = l >= j && l < n ? Rj[l] : c128{};
      }
      const c128 pjj = Pj[j], pj1 = Pj[j + 1], rjj = Rj[j], rj1 = Rj[j + 1];
      const c128 S{wsum(ps.re), wsum(ps.im)};                 // p^H v of column j - 1
      aw = cscale(-0.5, cmul(tau_prev, S));                   // w = p + aw v
      const c128 vx{vnx_re, vnx_im};
      wj = pjj + aw;                                          // v_{j-1}[j] = 1
#pragma unroll
      for (int c = 0; c < kHCols; ++c) {
        wr[c] = Pl[c] + cmul(aw, vr[c]);
        R[c] = R[c] - (conj(wr[c]) + cmul(wj, conj(vr[c])));
      }
      dj = rjj.re - 2.0 * wj.re;
      const c128 w1 = pj1 + cmul(aw, vx);
      alpha = conj(rj1 - (conj(w1) + cmul(wj, conj(vx))));
    } else {
#pragma unroll
      for (int c = 0; c < kHCols; ++c) {
        const int l = t + kHThreads * c;
        R[c] = l < n ? Rj[l] : c128{};
        wr[c] = c128{};
      }
      dj = Rj[0].re;
      alpha = conj(Rj[1]);
    }
    // the reflector of column j: x = conj(row j) right of the subdiagonal
    double sq = 0.0;
#pragma unroll
    for (int c = 0; c < kHCols; ++c) {
      const int l = t + kHThreads * c;
      if (l >= j + 2 && l < n) sq += cabs2(R[c]);
    }
    double sigma;
    {
      double v = wsum(sq);
      double* rb = red + kHWaves * par;
      par ^= 1;
      if (lane == 0) rb[wv] = v;
      __syncthreads();
      sigma = 0.0;
#pragma unroll
      for (int w = 0; w < kHWaves; ++w) sigma += rb[w];
    }
    double beta = alpha.re;
    c128 tj{}, scale{};
    if (sigma > 0.0 || alpha.im != 0.0) {
      beta = -copysign(sqrt(cabs2(alpha) + sigma), alpha.re);
      tj = c128{(beta - alpha.re) / beta, -alpha.im / beta};
      const c128 den{alpha.re - beta, alpha.im};              // 1 / (alpha - beta)
      const double q = 1.0 / cabs2(den);
      scale = c128{den.re * q, -den.im * q};
    }
    c128 vn[kHCols];
#pragma unroll
    for (int c = 0; c < kHCols; ++c) {
      const int l = t + kHThreads * c;
      vn[c] = l == j + 1 ? c128{1.0, 0.0}
                         : (l >= j + 2 && l < n ? cmul(conj(R[c]), scale) : c128{});
      if (l < n && l >= g && (l - g) % G == 0) {
        const int rl = (l - g) / G;
        on_re[rl] = vn[c].re;
        on_im[rl] = vn[c].im;
        Vr[(long)j * n + l] = vn[c];
      }
      if (l == j + 2) {
        vnx_re = vn[c].re;
        vnx_im = vn[c].im;
      }
    }
    if (g == j % G && t == 0) {
      dd[j] = dj;
      ee[j] = beta;
      tt[j] = tj;
    }
    if (j == n - 2) {   // the last diagonal gets column j-1's update
      if (g == (n - 1) % G) {
        const int rl = (n - 1 - g) / G;
        const c128 vi{ov_re[rl], ov_im[rl]};
        const c128 wi = c128{op_re[rl], op_im[rl]} + cmul(aw, vi);
#pragma unroll
        for (int r = 0; r < kHRows; ++r)
#pragma unroll
          for (int c = 0; c < kHCols; ++c)
            if (r == rl && t + kHThreads * c == n - 1)
              dd[n - 1] = x[r][c].re - 2.0 * cmul(vi, conj(wi)).re;
      }
      break;
    }
    // one pass over the own rows: update by (v_{j-1}, w_{j-1}), p_j = tau_j A v_j on rows > j,
    // row j+1 published by its owner
    const int rpub = (j + 1 < n - 1 && g == (j + 1) % G) ? (j + 1 - g) / G : -1;
    double (*rb)[2 * kHRows] = red8[par8];
    par8 ^= 1;
    double s_re[kHRows], s_im[kHRows];
#pragma unroll
    for (int r = 0; r < kHRows; ++r) {
      const c128 vi{ov_re[r], ov_im[r]};
      const c128 wi = c128{op_re[r], op_im[r]} + cmul(aw, vi);
      const bool live = r < nr && g + r * G >= j;
      c128 acc{};
#pragma unroll
      for (int c = 0; c < kHCols; ++c) {
        const int l = t + kHThreads * c;
        if (live && l >= j && l < n)
          x[r][c] = x[r][c] - (cmul(vi, conj(wr[c])) + cmul(wi, conj(vr[c])));
        acc = acc + cmul(x[r][c], vn[c]);
        if (r == rpub && l > j && l < n) st_shared_c(Rp + (long)(j + 1) * a.ns + l, x[r][c]);
      }
      const bool keep = r < nr && g + r * G > j;
      s_re[r] = keep ? acc.re : 0.0;
      s_im[r] = keep ? acc.im : 0.0;
    }
    wsum8_store(s_re, rb[wv], lane);
    wsum8_store(s_im, rb[wv] + kHRows, lane);
    __syncthreads();
    c128 partial{};
    if (t < nr) {
      const int i = g + t * G;
      c128 pv{};
      if (i >= j + 1) {
        c128 sum{};
#pragma unroll
        for (int w = 0; w < kHWaves; ++w) sum = sum + c128{rb[w][t], rb[w][kHRows + t]};
        pv = cmul(tj, sum);
        st_shared_c(P + (long)j * a.ns + i, pv);
        partial = cmul(conj(pv), c128{on_re[t], on_im[t]});
      }
      op_re[t] = pv.re;
      op_im[t] = pv.im;
      ov_re[t] = on_re[t];
      ov_im[t] = on_im[t];
    }
    if (wv == 0) {
      const c128 ps{wsum(partial.re), wsum(partial.im)};
      if (t == 0) st_shared_c(part + (long)j * a.gs + g, ps);
    }
    tau_prev = tj;
#pragma unroll
    for (int c = 0; c < kHCols; ++c) vr[c] = vn[c];
    if (!grid_sync_of(bar, epoch++, &flag, G, g)) return;
  }
}

int htrd_grid(int n) { return (n + kHRows - 1) / kHRows; }

// Threads per workgroup of the reduction: one thread per two columns, so orders up to 512 take
// 256 threads, and twice as many matrices stay resident per launch.
int htrd_threads(int n) { return n <= 512 ? 256 : 512; }

const void* htrd_kernel(int n) {
  return n <= 512 ? reinterpret_cast<const void*>(&hetrd_coop_kernel<256>)
                  : reinterpret_cast<const void*>(&hetrd_coop_kernel<512>);
}

// Per matrix: P and Rp slots (n x ns complex), partials (n x gs complex), barrier words.
size_t htrd_ws_per_matrix(int n) {
  const int G = htrd_grid(n), ns = trd_slot(n), gs = (G + 31) / 32 * 32 + 32;
  return align_up((size_t)n * ns * 16, 256) * 2 + align_up((size_t)n * gs * 16, 256) +
         align_up(kBarWords * sizeof(unsigned), 256);
}

// Complex back-transformation helpers.
// Vc[i][j] = Vr[j][i] and Vh[i][j] = conj(Vr[j][i]) for reflectors j < n - 1 (zero otherwise),
// for matrix blockIdx.z (Vr at z n n, Vc and Vh at z n NV).
__global__ void transpose_c_kernel(int n, int NV, const c128* __restrict__ src,
                                   c128* __restrict__ Vc, c128* __restrict__ Vh) {
  __shared__ double tre[32][33], tim[32][33];
  src += blockIdx.z * (long)n * n;          // matrix blockIdx.z
  Vc += blockIdx.z * (long)n * NV;
  Vh += blockIdx.z * (long)n * NV;
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int j = by + r, i = bx + threadIdx.x;
    const c128 v = (j < n - 1 && i < n) ? src[(long)j * n + i] : c128{};
    tre[r][threadIdx.x] = v.re;
    tim[r][threadIdx.x] = v.im;
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int i = bx + r, j = by + threadIdx.x;
    if (i < n && j < NV) {
      const c128 v{tre[threadIdx.x][r], tim[threadIdx.x][r]};
      Vc[(long)i * NV + j] = v;
      Vh[(long)i * NV + j] = conj(v);
    }
  }
}

// Panel blockIdx.x of matrix blockIdx.y: T of Q_b = I - V T V^H (T[0:c, c] = -tau_c T[0:c, 0:c]
// S[0:c, c], T[c][c] = tau_c with S = V^H V), written as A_b[p][q] = -T[q][p], the K-major left
// operand of Y' = -T Y. S, tau and negTt of matrix y at y sS, y stau, y sS.
__global__ void __launch_bounds__(256) larft_c_kernel(int nrefl, const c128* S, const c128* tau,
                                                      c128* negTt, long sS, long stau) {
  __shared__ double Tre[kBtNb][kBtNb + 1], Tim[kBtNb][kBtNb + 1];
  __shared__ double mre[kBtNb], mim[kBtNb];
  const int b = blockIdx.x, jb = b * kBtNb, nb = min(kBtNb, nrefl - jb), t = threadIdx.x;
  S += blockIdx.y * sS;
  tau += blockIdx.y * stau;
  negTt += blockIdx.y * sS;
  const c128* Sb = S + (long)b * kBtNb * kBtNb;
  for (int idx = t; idx < kBtNb * kBtNb; idx += blockDim.x) {
    Tre[idx / kBtNb][idx % kBtNb] = 0.0;
    Tim[idx / kBtNb][idx % kBtNb] = 0.0;
  }
  __syncthreads();
  for (int c = 0; c < nb; ++c) {
    const c128 tc = tau[jb + c];
    if (t < c) {
      c128 s{};
      for (int q = t; q < c; ++q) s = s + cmul(c128{Tre[t][q], Tim[t][q]}, Sb[q * kBtNb + c]);
      mre[t] = s.re;
      mim[t] = s.im;
    }
    __syncthreads();
    if (t < c) {
      const c128 v = cmul(tc, c128{mre[t], mim[t]});
      Tre[t][c] = -v.re;
      Tim[t][c] = -v.im;
    }
    if (t == 0) {
      Tre[c][c] = tc.re;
      Tim[c][c] = tc.im;
    }
    __syncthreads();
  }
  c128* out = negTt + (long)b * kBtNb * kBtNb;
  for (int idx = t; idx < kBtNb * kBtNb; idx += blockDim.x) {
    const int p = idx / kBtNb, q = idx % kBtNb;
    out[idx] = c128{-Tre[q][p], -Tim[q][p]};
  }
}

// Zc (n x k complex, matrix blockIdx.y at y cnt) = Zw (n x k real, matrix y at y sw).
__global__ void real_to_c_kernel(long cnt, const double* __restrict__ src, long sw,
                                 c128* __restrict__ dst) {
  src += blockIdx.y * sw;
  dst += blockIdx.y * cnt;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < cnt; i += (long)gridDim.x * 256)
    dst[i] = c128{src[i], 0.0};
}

__global__ void heev_status_kernel(const int* err, const int* sweeps, int nleaf, int* info) {
  int bad = 0;
  for (int i = threadIdx.x; i < nleaf; i += blockDim.x) bad |= sweeps[i] < 0;
  bad = __syncthreads_or(bad);
  if (threadIdx.x == 0) *info |= (*err ? 1 : 0) | (bad ? 2 : 0);
}

int gemm_c(int M, int N, int K, const c128* A, long lda, const c128* B, long ldb, c128* C,
           long ldc, int batch, long sA, long sB, long sC, int acc, hipStream_t st) {
  return tnh_gemm_tn(TNH_C128, M, N, K, A, lda, B, ldb, C, ldc, batch, sA, sB, sC, acc, 0, st);
}

struct HeevPlan {
  SyevPlan sp;
  int batch, mpl;   // matrices, matrices per reduction launch
};

// Workspace pieces of the batched complex solve, in carving order: the real solver's pieces for
// the divide and conquer of all h.batch tridiagonals at once, the reduction's exchange slots and
// barrier words for one launch of h.mpl matrices, then per matrix the reflectors, T and the
// back-transformation's blocks.
template <typename F>
void heev_pieces(const HeevPlan& h, F&& f) {
  const SyevPlan& p = h.sp;
  const size_t n = p.n, k = p.k, NV = (size_t)p.npan * kBtNb, B = h.batch, C = h.mpl;
  const size_t G = htrd_grid(p.n), ns = trd_slot(p.n), gs = (G + 31) / 32 * 32 + 32;
  f("syev", syev_ws_bytes(p));
  f("P", C * n * ns * 16);
  f("Rp", C * n * ns * 16);
  f("part", C * n * gs * 16);
  f("bars", C * kBarWords * sizeof(unsigned));
  f("err", 256);
  f("Vr", B * n * n * 16);
  f("d", B * n * 8);
  f("e", B * n * 8);
  f("tau", B * n * 16);
  f("Vc", B * n * NV * 16);
  f("Vh", B * n * NV * 16);
  f("S", B * p.npan * kBtNb * kBtNb * 16);
  f("negTt", B * p.npan * kBtNb * kBtNb * 16);
  f("Zc", B * n * k * 16);
  f("Y", B * kBtNb * k * 16);
  f("Y2", B * kBtNb * k * 16);
}

size_t heev_ws_bytes(const HeevPlan& h) {
  size_t tot = 0;
  heev_pieces(h, [&](const char*, size_t b) { tot += align_up(b, 256); });
  return tot + 256;
}

// Matrices per reduction launch: as many whole matrices as the device keeps resident at once.
int heev_mpl(int n, int batch, int* mpl) {
  int dev = 0, cus = 0, per_cu = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, htrd_kernel(n), htrd_threads(n),
                                                   0) != hipSuccess) {
    set_error("heev: device query failed");
    return TNH_ERR_HIP;
  }
  const int G = htrd_grid(n), cap = (cus * per_cu) / G;
  TNH_REQUIRE(cap >= 1, TNH_ERR_UNSUPPORTED,
              "heev: the %d workgroups of one matrix are not resident at once", G);
  *mpl = std::min(cap, batch);
  return TNH_OK;
}

// The reductions, h.mpl matrices per cooperative launch; one divide and conquer over all the
// tridiagonals; Z = Q Z_T for all matrices at once (batched GEMMs per block of reflectors).
int heev_top(const HeevPlan& h, const c128* A, long lda, long sA, double* w, long sw, c128* Z,
             long ldz, long sZ, void* ws, int* info, hipStream_t st) {
  const SyevPlan& p = h.sp;
  std::vector<char*> ptrs;
  {
    char* base = reinterpret_cast<char*>(align_up(reinterpret_cast<size_t>(ws), 256));
    size_t off = 0;
    heev_pieces(h, [&](const char*, size_t b) {
      ptrs.push_back(base + off);
      off += align_up(b, 256);
    });
  }
  int q = 0;
  auto cp = [&]() { return reinterpret_cast<c128*>(ptrs[q++]); };
  auto dp = [&]() { return reinterpret_cast<double*>(ptrs[q++]); };
  char* ws_syev = ptrs[q++];
  c128* P = cp();
  c128* Rp = cp();
  c128* part = cp();
  unsigned* bars = reinterpret_cast<unsigned*>(ptrs[q++]);
  int* err = reinterpret_cast<int*>(ptrs[q++]);
  c128* Vr = cp();
  double* d = dp();
  double* e = dp();
  c128* tau = cp();
  c128* Vc = cp();
  c128* Vh = cp();
  c128* S = cp();
  c128* negTt = cp();
  c128* Zc = cp();
  c128* Y = cp();
  c128* Y2 = cp();
  const int n = p.n, k = p.k, B = h.batch, G = htrd_grid(n), ns = trd_slot(n);
  const int gs = (G + 31) / 32 * 32 + 32;
  const long nn = (long)n * n;
  const int NV = p.npan * kBtNb, nrefl = n - 1;
  SyevBufs b = syev_carve(p, ws_syev);

  hipError_t he = hipMemsetAsync(err, 0, sizeof(int), st);
  if (he == hipSuccess && info) he = hipMemsetAsync(info, 0, sizeof(int), st);
  TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP, "heev: reset: %s", hipGetErrorString(he));
  for (int m0 = 0; m0 < B; m0 += h.mpl) {
    const int cnt = std::min(h.mpl, B - m0);
    he = hipMemsetAsync(bars, 0, (size_t)cnt * kBarWords * sizeof(unsigned), st);
    TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP, "heev: barrier reset: %s", hipGetErrorString(he));
    HtrdArgs args{n,  ns, gs, G, A + (long)m0 * sA, lda, sA, Vr + m0 * nn, d + (long)m0 * n,
                  e + (long)m0 * n, tau + (long)m0 * n, P, Rp, part, bars, err};
    void* kargs[] = {&args};
    he = hipLaunchCooperativeKernel(htrd_kernel(n), dim3(G * cnt), dim3(htrd_threads(n)), kargs,
                                    0u, st);
    TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP,
                "heev: cooperative launch of %d workgroups refused: %s", G * cnt,
                hipGetErrorString(he));
  }
  int rc = run_dc(p, b, d, n, e, n, w, sw, st);
  if (rc != TNH_OK) return rc;

  // Z = Q Z_T for all matrices: Zc = Z_T, then the blocks of reflectors last first
  const bool direct = ldz == k && sZ == (long)n * k;   // Z itself holds the running product
  c128* Zr = direct ? Z : Zc;
  const long cntz = (long)n * k, sVc = (long)n * NV, sT = (long)p.npan * kBtNb * kBtNb;
  const long sY = (long)kBtNb * k;
  hipLaunchKernelGGL(real_to_c_kernel, dim3((unsigned)std::min<long>((cntz + 255) / 256, 1024), B),
                     dim3(256), 0, st, cntz, b.Zw, (long)p.N * k, Zr);
  hipLaunchKernelGGL(transpose_c_kernel, dim3((n + 31) / 32, (NV + 31) / 32, B), dim3(32, 8), 0,
                     st, n, NV, Vr, Vc, Vh);
  TNH_CHECK_LAUNCH("heev transpose");
  for (int m = 0; m < B; ++m) {
    rc = gemm_c(kBtNb, kBtNb, n, Vh + m * sVc, NV, Vc + m * sVc, NV, S + m * sT, kBtNb, p.npan,
                kBtNb, kBtNb, (long)kBtNb * kBtNb, 0, st);
    if (rc != TNH_OK) return rc;
  }
  hipLaunchKernelGGL(larft_c_kernel, dim3(p.npan, B), dim3(256), 0, st, nrefl, S, tau, negTt, sT,
                     (long)n);
  TNH_CHECK_LAUNCH("heev larft");
  for (int bb = p.npan - 1; bb >= 0; --bb) {
    const int jb = bb * kBtNb, nb = std::min(kBtNb, nrefl - jb), rows = n - jb;
    rc = gemm_c(nb, k, rows, Vh + (long)jb * NV + jb, NV, Zr + (long)jb * k, k, Y, k, B, sVc,
                cntz, sY, 0, st);
    if (rc != TNH_OK) return rc;
    rc = gemm_c(nb, k, nb, negTt + (long)bb * kBtNb * kBtNb, kBtNb, Y, k, Y2, k, B, sT, sY, sY, 0,
                st);
    if (rc != TNH_OK) return rc;
    rc = gemm_c(rows, k, nb, Vr + (long)jb * n + jb, n, Y2, k, Zr + (long)jb * k, k, B, nn, sY,
                cntz, 1, st);
    if (rc != TNH_OK) return rc;
  }
  if (!direct)
    for (int m = 0; m < B; ++m) {
      he = hipMemcpy2DAsync(Z + m * sZ, (size_t)ldz * 16, Zc + m * cntz, (size_t)k * 16,
                            (size_t)k * 16, (size_t)n, hipMemcpyDeviceToDevice, st);
      TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP, "heev: vector copy: %s", hipGetErrorString(he));
    }
  if (info) {
    hipLaunchKernelGGL(heev_status_kernel, dim3(1), dim3(256), 0, st, err, b.sweeps,
                       B * p.N / p.s, info);
    TNH_CHECK_LAUNCH("heev status");
  }
  return TNH_OK;
}

// ---- thin QR of tall matrices (float64 and complex128) -------------------------------------------
// A = Q R (m x n, n <= m <= 2048, n <= 2048) with LAPACK's geqrf reflectors (H_j = I - tau_j v_j
// v_j^H, v_j[j] = 1, H_j^H (alpha; x) = (beta; 0), beta real, so R has a real diagonal and is
// the R that geqrf returns), and Q = H_0 ... H_{n-1} [I; 0] (ungqr's Q) by WY blocks. The
// reduction is one cooperative launch per group of resident matrices: workgroup g of a matrix
// holds columns g, g + G, ... of A (as rows of A^H, conjugated) in registers, 8 of them, each
// thread two of the m entries. Column j: its owner forms the reflector from its register copy
// and publishes v_j (one grid barrier), then every workgroup applies H_j^H to its own columns
// right of j (a dot product with v_j and an update per column, reduced within the workgroup),
// and the owner of column j+1 forms the next reflector from its updated copy at once.
__device__ __forceinline__ double sre(double x) { return x; }
__device__ __forceinline__ double sre(c128 x) { return x.re; }
__device__ __forceinline__ double sim(double) { return 0.0; }
__device__ __forceinline__ double sim(c128 x) { return x.im; }
__device__ __forceinline__ double cj(double x) { return x; }
__device__ __forceinline__ c128 cj(c128 x) { return conj(x); }
__device__ __forceinline__ double nrm2(double x) { return x * x; }
__device__ __forceinline__ double nrm2(c128 x) { return cabs2(x); }
template <typename T>
__device__ __forceinline__ T mk(double re, double im);
template <>
__device__ __forceinline__ double mk<double>(double re, double) { return re; }
template <>
__device__ __forceinline__ c128 mk<c128>(double re, double im) { return c128{re, im}; }
__device__ __forceinline__ void st_shared_s(double* p, double v) { st_shared(p, v); }
__device__ __forceinline__ void st_shared_s(c128* p, c128 v) { st_shared_c(p, v); }

constexpr int kQrRows = 8;   // columns of A per workgroup

template <typename T>
struct QrArgs {
  int m, n, G;           // A is m x n; G workgroups per matrix
  const T* A;            // matrix q of the launch at A + q sA, row stride lda
  long lda, sA;
  T* V;                  // per matrix n slots of ldv: v_j (zero above j, 1 at j), tau_j at m
  int ldv;               // slot length: m + 1 rounded up to 32 elements (whole cache lines)
  T* tau;                // per matrix n: tau_j again, for the host side
  T* R;                  // per matrix n x n at R + q sR, row stride ldr (upper; zeros below)
  long ldr, sR;
  unsigned* bar;         // per matrix kBarWords
  int* err;
};

template <typename T, int NT, int C>
__global__ void __launch_bounds__(NT) geqrf_coop_kernel(QrArgs<T> a) {
  constexpr int NW = NT / 64, RW = kQrRows;
  constexpr bool kCplx = sizeof(T) == 16;
  __shared__ double red[2 * NW];
  __shared__ double red8[NW][2 * RW];
  __shared__ double dots[2 * RW];
  __shared__ double al_re, al_im;
  __shared__ int flag;
  const int m = a.m, n = a.n, G = a.G, mb = blockIdx.x / G, g = blockIdx.x % G, t = threadIdx.x;
  const int lane = t & 63, wv = t >> 6;
  const int nr = (n - g + G - 1) / G;
  const T* A = a.A + mb * a.sA;
  const int ldv = a.ldv;
  T* V = a.V + (long)mb * n * ldv;
  T* tau = a.tau + (long)mb * n;
  T* R = a.R + mb * a.sR;
  const GridBar bar{a.bar + (size_t)mb * kBarWords, a.err};
  int par = 0;

  T x[RW][C];   // x[q][c] = conj(A[l][g + q G]) at l = t + NT c
#pragma unroll
  for (int q = 0; q < RW; ++q)
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int l = t + NT * c;
      x[q][c] = (q < nr && l < m) ? cj(A[(long)l * a.lda + g + q * G]) : T{};
    }

  // the owner of column p forms reflector p from its register copy, publishes v_p and tau_p,
  // writes column p of R and leaves (beta, 0, ...) in its registers
  auto pivot = [&](int p) {
    const int qp = (p - g) / G;
    double sq = 0.0;
#pragma unroll
    for (int q = 0; q < RW; ++q)
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int l = t + NT * c;
        if (q == qp && l > p && l < m) sq += nrm2(x[q][c]);
        if (q == qp && l == p) {
          al_re = sre(x[q][c]);
          al_im = -sim(x[q][c]);
        }
      }
    double v = wsum(sq);
    double* rb = red + NW * par;
    par ^= 1;
    if (lane == 0) rb[wv] = v;
    __syncthreads();
    double sigma = 0.0;
#pragma unroll
    for (int w = 0; w < NW; ++w) sigma += rb[w];
    const double ar = al_re, ai = al_im;
    double beta = ar;
    T tj{}, scale{};
    if (sigma > 0.0 || ai != 0.0) {
      beta = -copysign(sqrt(ar * ar + ai * ai + sigma), ar);
      tj = mk<T>((beta - ar) / beta, -ai / beta);
      const double dr = ar - beta, di = ai, qq = 1.0 / (dr * dr + di * di);
      scale = mk<T>(dr * qq, -di * qq);
    }
#pragma unroll
    for (int q = 0; q < RW; ++q)
#pragma unroll
      for (int c = 0; c < C; ++c) {
        const int l = t + NT * c;
        if (q == qp) {
          const T vv = l == p ? mk<T>(1.0, 0.0) : (l > p && l < m ? cj(x[q][c]) * scale : T{});
          if (l < m) st_shared_s(V + (long)p * ldv + l, vv);
          if (l < n) R[(long)l * a.ldr + p] = l < p ? cj(x[q][c]) : (l == p ? mk<T>(beta, 0.0) : T{});
          x[q][c] = l == p ? mk<T>(beta, 0.0) : (l > p ? T{} : x[q][c]);
        }
      }
    if (t == 0) {
      st_shared_s(V + (long)p * ldv + m, tj);
      tau[p] = tj;
    }
  };

  if (g == 0) pivot(0);
  if (n == 1) return;
  unsigned epoch = 1;
  if (!grid_sync_of(bar, epoch++, &flag, G, g)) return;
  for (int j = 0; j < n - 1; ++j) {
    // apply H_j^H to the own columns right of j: x -= tau_j (x . v_j) conj(v_j)
    T vj[C];
#pragma unroll
    for (int c = 0; c < C; ++c) {
      const int l = t + NT * c;
      vj[c] = (l >= j && l < m) ? V[(long)j * ldv + l] : T{};
    }
    const T tj = V[(long)j * ldv + m];
    double s_re[RW], s_im[RW];
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      T acc{};
#pragma unroll
      for (int c = 0; c < C; ++c) acc = acc + x[q][c] * vj[c];
      const bool live = q < nr && g + q * G > j;
      s_re[q] = live ? sre(acc) : 0.0;
      s_im[q] = live ? sim(acc) : 0.0;
    }
    wsum8_store(s_re, red8[wv], lane);
    if constexpr (kCplx) wsum8_store(s_im, red8[wv] + RW, lane);
    __syncthreads();
    if (t < (kCplx ? 2 * RW : RW)) {
      double sum = 0.0;
#pragma unroll
      for (int w = 0; w < NW; ++w) sum += red8[w][t];
      dots[t] = sum;
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < RW; ++q) {
      const bool live = q < nr && g + q * G > j;
      const T f = tj * mk<T>(dots[q], kCplx ? dots[RW + q] : 0.0);
#pragma unroll
      for (int c = 0; c < C; ++c)
        if (live) x[q][c] = x[q][c] - f * cj(vj[c]);
    }
    if (g == (j + 1) % G) pivot(j + 1);
    if (j + 1 < n - 1 && !grid_sync_of(bar, epoch++, &flag, G, g)) return;
  }
}

// Vc[i][j] = V[j][i], Vh[i][j] = conj(V[j][i]) for reflectors j < nrefl (rows of length m),
// zero for nrefl <= j < NV (Vh may be null for real matrices); reflector j at src + j lds.
template <typename T>
__global__ void qr_transpose_kernel(int m, int nrefl, int NV, const T* __restrict__ src, int lds,
                                    long ssrc, T* __restrict__ Vc, T* __restrict__ Vh, long sdst) {
  __shared__ double tre[32][33], tim[32][33];
  src += blockIdx.z * ssrc;   // matrix blockIdx.z
  Vc += blockIdx.z * sdst;
  if (Vh) Vh += blockIdx.z * sdst;
  const int bx = blockIdx.x * 32, by = blockIdx.y * 32;
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int j = by + r, i = bx + threadIdx.x;
    const T v = (j < nrefl && i < m) ? src[(long)j * lds + i] : T{};
    tre[r][threadIdx.x] = sre(v);
    tim[r][threadIdx.x] = sim(v);
  }
  __syncthreads();
  for (int r = threadIdx.y; r < 32; r += blockDim.y) {
    const int i = bx + r, j = by + threadIdx.x;
    if (i < m && j < NV) {
      const double vr = tre[threadIdx.x][r], vi = tim[threadIdx.x][r];
      Vc[(long)i * NV + j] = mk<T>(vr, vi);
      if (Vh) Vh[(long)i * NV + j] = mk<T>(vr, -vi);
    }
  }
}

// Q (m x n, row stride ldq, cnt matrices at stride sQ) = [I; 0].
template <typename T>
__global__ void eye_kernel(int m, int n, T* Q, long ldq, long sQ) {
  T* Qm = Q + blockIdx.y * sQ;
  const long cnt = (long)m * n;
  for (long i = blockIdx.x * 256L + threadIdx.x; i < cnt; i += (long)gridDim.x * 256) {
    const int r = (int)(i / n), c = (int)(i % n);
    Qm[(long)r * ldq + c] = mk<T>(r == c ? 1.0 : 0.0, 0.0);
  }
}

__global__ void qr_status_kernel(const int* err, int* info) {
  if (threadIdx.x == 0) *info |= *err ? 1 : 0;
}

// Threads per workgroup for rows of length m: two entries per thread up to 1024, float64 rows up
// to 2048 as 512 threads of four (the tridiagonalisation's measured best shape); complex rows of
// more than 1024 entries would not fit the registers and are refused by the entry.
template <typename T>
int qr_threads(int m) {
  return m <= 512 ? 256 : 512;
}

template <typename T>
int qr_max_m() {
  return sizeof(T) == 16 ? 1024 : 2048;
}

template <typename T>
const void* qr_kernel(int m) {
  if (m <= 512) return reinterpret_cast<const void*>(&geqrf_coop_kernel<T, 256, 2>);
  if constexpr (sizeof(T) == 8) {
    if (m > 1024) return reinterpret_cast<const void*>(&geqrf_coop_kernel<T, 512, 4>);
  }
  return reinterpret_cast<const void*>(&geqrf_coop_kernel<T, 512, 2>);
}

int qr_grid(int n) { return (n + kQrRows - 1) / kQrRows; }

// Matrices per reduction launch (all of them resident at once).
template <typename T>
int qr_mpl(int m, int n, int batch, int* mpl) {
  int dev = 0, cus = 0, per_cu = 0;
  const int nt = qr_threads<T>(m);
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, qr_kernel<T>(m), nt, 0) !=
          hipSuccess) {
    set_error("qr: device query failed");
    return TNH_ERR_HIP;
  }
  const int cap = (cus * per_cu) / qr_grid(n);
  TNH_REQUIRE(cap >= 1, TNH_ERR_UNSUPPORTED,
              "qr: the %d workgroups of one matrix are not resident at once", qr_grid(n));
  *mpl = std::min(cap, batch);
  return TNH_OK;
}

int qr_ldv(int m) { return (m + 1 + 31) / 32 * 32; }

template <typename T, typename F>
void qr_pieces(int m, int n, int mpl, F&& f) {
  const size_t npan = (n + kBtNb - 1) / kBtNb, NV = npan * kBtNb, B = mpl, E = sizeof(T);
  f(B * n * qr_ldv(m) * E);                  // V slots
  f(B * n * E);                              // tau
  f(B * kBarWords * sizeof(unsigned));       // barrier words
  f(256);                                    // err
  f(B * m * NV * E);                         // Vc
  f(E == 16 ? B * m * NV * E : 0);           // Vh (complex only)
  f(B * npan * kBtNb * kBtNb * E);           // S
  f(B * npan * kBtNb * kBtNb * E);           // negTt
  f(B * kBtNb * n * E);                      // Y
  f(B * kBtNb * n * E);                      // Y2
}

template <typename T>
size_t qr_ws_bytes(int m, int n, int mpl) {
  size_t tot = 0;
  qr_pieces<T>(m, n, mpl, [&](size_t b) { tot += align_up(b, 256); });
  return tot + 256;
}

inline void larft_any(int np, int cnt, int nrefl, const double* S, const double* tau,
                      double* negTt, long sS, long stau, hipStream_t st) {
  hipLaunchKernelGGL(larft_kernel, dim3(np, cnt), dim3(256), 0, st, nrefl, S, tau, negTt, sS,
                     stau);
}
inline void larft_any(int np, int cnt, int nrefl, const c128* S, const c128* tau, c128* negTt,
                      long sS, long stau, hipStream_t st) {
  hipLaunchKernelGGL(larft_c_kernel, dim3(np, cnt), dim3(256), 0, st, nrefl, S, tau, negTt, sS,
                     stau);
}

template <typename T>
int qr_run(int dtype, int batch, int m, int n, int mpl, const T* A, long lda, long sA, T* Q,
           long ldq, long sQ, T* R, long ldr, long sR, int* info, void* ws, hipStream_t st) {
  std::vector<char*> ptrs;
  {
    char* base = reinterpret_cast<char*>(align_up(reinterpret_cast<size_t>(ws), 256));
    size_t off = 0;
    qr_pieces<T>(m, n, mpl, [&](size_t b) {
      ptrs.push_back(base + off);
      off += align_up(b, 256);
    });
  }
  T* V = reinterpret_cast<T*>(ptrs[0]);
  T* tau = reinterpret_cast<T*>(ptrs[1]);
  unsigned* bars = reinterpret_cast<unsigned*>(ptrs[2]);
  int* err = reinterpret_cast<int*>(ptrs[3]);
  T* Vc = reinterpret_cast<T*>(ptrs[4]);
  T* Vh = sizeof(T) == 16 ? reinterpret_cast<T*>(ptrs[5]) : Vc;
  T* S = reinterpret_cast<T*>(ptrs[6]);
  T* negTt = reinterpret_cast<T*>(ptrs[7]);
  T* Y = reinterpret_cast<T*>(ptrs[8]);
  T* Y2 = reinterpret_cast<T*>(ptrs[9]);
  const int G = qr_grid(n), nt = qr_threads<T>(m), npan = (n + kBtNb - 1) / kBtNb;
  const int NV = npan * kBtNb, ldv = qr_ldv(m);
  const long sV = (long)n * ldv, sVc = (long)m * NV, sT = (long)npan * kBtNb * kBtNb,
             sY = (long)kBtNb * n;
  hipError_t he = hipMemsetAsync(err, 0, sizeof(int), st);
  if (he == hipSuccess && info) he = hipMemsetAsync(info, 0, sizeof(int), st);
  TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP, "qr: reset: %s", hipGetErrorString(he));
  for (int m0 = 0; m0 < batch; m0 += mpl) {
    const int cnt = std::min(mpl, batch - m0);
    he = hipMemsetAsync(bars, 0, (size_t)cnt * kBarWords * sizeof(unsigned), st);
    TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP, "qr: barrier reset: %s", hipGetErrorString(he));
    QrArgs<T> args{m, n, G, A + (long)m0 * sA, lda, sA, V, ldv, tau, R + (long)m0 * sR, ldr, sR,
                   bars, err};
    void* kargs[] = {&args};
    he = hipLaunchCooperativeKernel(qr_kernel<T>(m), dim3(G * cnt), dim3(nt), kargs, 0u, st);
    TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP, "qr: cooperative launch of %d workgroups: %s",
                G * cnt, hipGetErrorString(he));
    T* Qm = Q + (long)m0 * sQ;
    hipLaunchKernelGGL(eye_kernel<T>, dim3(std::min(((long)m * n + 255) / 256, 1024L), cnt),
                       dim3(256), 0, st, m, n, Qm, ldq, sQ);
    TNH_CHECK_LAUNCH("qr eye");
    hipLaunchKernelGGL(qr_transpose_kernel<T>, dim3((m + 31) / 32, (NV + 31) / 32, cnt),
                       dim3(32, 8), 0, st, m, n, NV, V, ldv, sV, Vc,
                       sizeof(T) == 16 ? Vh : nullptr, sVc);
    TNH_CHECK_LAUNCH("qr transpose");
    for (int q = 0; q < cnt; ++q) {
      int rc = tnh_gemm_tn(dtype, kBtNb, kBtNb, m, Vh + q * sVc, NV, Vc + q * sVc, NV, S + q * sT,
                           kBtNb, npan, kBtNb, kBtNb, (long)kBtNb * kBtNb, 0, 0, st);
      if (rc != TNH_OK) return rc;
    }
    larft_any(npan, cnt, n, S, tau, negTt, sT, (long)n, st);
    TNH_CHECK_LAUNCH("qr larft");
    // Q <- H_0 ... H_{n-1} [I; 0], last block first; block b changes only columns >= jb
    for (int bb = npan - 1; bb >= 0; --bb) {
      const int jb = bb * kBtNb, nb = std::min(kBtNb, n - jb), rows = m - jb, nc = n - jb;
      int rc = tnh_gemm_tn(dtype, nb, nc, rows, Vh + (long)jb * NV + jb, NV,
                           Qm + (long)jb * ldq + jb, ldq, Y, nc, cnt, sVc, sQ, sY, 0, 0, st);
      if (rc != TNH_OK) return rc;
      rc = tnh_gemm_tn(dtype, nb, nc, nb, negTt + (long)bb * kBtNb * kBtNb, kBtNb, Y, nc, Y2, nc,
                       cnt, sT, sY, sY, 0, 0, st);
      if (rc != TNH_OK) return rc;
      rc = tnh_gemm_tn(dtype, rows, nc, nb, V + (long)jb * ldv + jb, ldv, Y2, nc,
                       Qm + (long)jb * ldq + jb, ldq, cnt, sV, sY, sQ, 1, 0, st);
      if (rc != TNH_OK) return rc;
    }
  }
  if (info) {
    hipLaunchKernelGGL(qr_status_kernel, dim3(1), dim3(64), 0, st, err, info);
    TNH_CHECK_LAUNCH("qr status");
  }
  return TNH_OK;
}

}  // namespace tnh

extern "C" size_t tnh_sytrd_ws_bytes(int n) {
  return n >= 1 && n <= tnh::kSyevMaxN ? tnh::trd_ws_bytes(n) : 0;
}

extern "C" int tnh_sytrd(int n, const double* A, long lda, double* d, double* e, double* tau,
                         double* Vr, int* status, void* ws, size_t ws_bytes, tnh_stream_t stream) {
  TNH_REQUIRE(n >= 1 && n <= tnh::kSyevMaxN, TNH_ERR_UNSUPPORTED,
              "tnh_sytrd: n = %d outside 1..%d", n, tnh::kSyevMaxN);
  TNH_REQUIRE(lda >= n && A && d && Vr && (n == 1 || (e && tau)), TNH_ERR_ARG,
              "tnh_sytrd: bad arguments (n=%d lda=%ld)", n, lda);
  TNH_REQUIRE(ws_bytes >= tnh::trd_ws_bytes(n), TNH_ERR_WS,
              "tnh_sytrd: workspace %zu < %zu bytes", ws_bytes, tnh::trd_ws_bytes(n));
  hipStream_t st = static_cast<hipStream_t>(stream);
  tnh::Arena ar(ws, ws_bytes);
  int* err = nullptr;
  int rc = tnh::launch_sytrd(n, A, lda, Vr, d, e, tau, ar, &err, st);
  if (rc != TNH_OK) return rc;
  if (status) {
    hipError_t he = hipMemcpyAsync(status, err, sizeof(int), hipMemcpyDeviceToDevice, st);
    TNH_REQUIRE(he == hipSuccess, TNH_ERR_HIP, "tnh_sytrd: status copy: %s",
                hipGetErrorString(he));
  }
  return TNH_OK;
}


extern "C" size_t tnh_syev_top_ws_bytes(int n, int k) {
  if (n < 2 || n > tnh::kSyevMaxN || k < 1 || k > n) return 0;
  return tnh::syev_ws_bytes(tnh::syev_plan(n, k));
}

extern "C" int tnh_syev_top(int n, int k, const double* A, long lda, double* w, double* Z,
                            long ldz, int* info, void* ws, size_t ws_bytes, tnh_stream_t stream) {
  TNH_REQUIRE(n >= 2 && n <= tnh::kSyevMaxN, TNH_ERR_UNSUPPORTED,
              "tnh_syev_top: n = %d outside 2..%d", n, tnh::kSyevMaxN);
  TNH_REQUIRE(k >= 1 && k <= n && lda >= n && ldz >= k && A && w && Z, TNH_ERR_ARG,
              "tnh_syev_top: bad arguments (n=%d k=%d lda=%ld ldz=%ld)", n, k, lda, ldz);
  const tnh::SyevPlan p = tnh::syev_plan(n, k);
  TNH_REQUIRE(ws && ws_bytes >= tnh::syev_ws_bytes(p), TNH_ERR_WS,
              "tnh_syev_top: workspace %zu < %zu bytes", ws_bytes, tnh::syev_ws_bytes(p));
  return tnh::syev_top(p, A, lda, w, Z, ldz, ws, ws_bytes, info,
                       static_cast<hipStream_t>(stream));
}

extern "C" size_t tnh_heev_top_ws_bytes(int batch, int n, int k) {
  if (batch < 1 || n < 2 || n > tnh::kHeevMaxN || k < 1 || k > n) return 0;
  int mpl = 1;
  if (tnh::heev_mpl(n, batch, &mpl) != TNH_OK) return 0;
  return tnh::heev_ws_bytes(tnh::HeevPlan{tnh::syev_plan(n, k, batch), batch, mpl});
}

extern "C" int tnh_heev_top(int batch, int n, int k, const void* A, long lda, long sA, double* w,
                            long sw, void* Z, long ldz, long sZ, int* info, void* ws,
                            size_t ws_bytes, tnh_stream_t stream) {
  TNH_REQUIRE(n >= 2 && n <= tnh::kHeevMaxN, TNH_ERR_UNSUPPORTED,
              "tnh_heev_top: n = %d outside 2..%d", n, tnh::kHeevMaxN);
  TNH_REQUIRE(batch >= 1 && k >= 1 && k <= n && lda >= n && ldz >= k && A && w && Z &&
                  (batch == 1 || (sA >= lda * (n - 1) + n && sw >= n && sZ >= ldz * (n - 1) + k)),
              TNH_ERR_ARG, "tnh_heev_top: bad arguments (batch=%d n=%d k=%d lda=%ld ldz=%ld)",
              batch, n, k, lda, ldz);
  int mpl = 1;
  int rc = tnh::heev_mpl(n, batch, &mpl);
  if (rc != TNH_OK) return rc;
  const tnh::HeevPlan h{tnh::syev_plan(n, k, batch), batch, mpl};
  TNH_REQUIRE(ws && ws_bytes >= tnh::heev_ws_bytes(h), TNH_ERR_WS,
              "tnh_heev_top: workspace %zu < %zu bytes", ws_bytes, tnh::heev_ws_bytes(h));
  return tnh::heev_top(h, static_cast<const tnh::c128*>(A), lda, sA, w, sw,
                       static_cast<tnh::c128*>(Z), ldz, sZ, ws, info,
                       static_cast<hipStream_t>(stream));
}

extern "C" size_t tnh_qr_ws_bytes(int dtype, int batch, int m, int n) {
  if (batch < 1 || n < 1 || n > m) return 0;
  int mpl = 1;
  if (dtype == TNH_F64) {
    if (m > tnh::qr_max_m<double>() || tnh::qr_mpl<double>(m, n, batch, &mpl) != TNH_OK) return 0;
    return tnh::qr_ws_bytes<double>(m, n, mpl);
  }
  if (dtype == TNH_C128) {
    if (m > tnh::qr_max_m<tnh::c128>() || tnh::qr_mpl<tnh::c128>(m, n, batch, &mpl) != TNH_OK)
      return 0;
    return tnh::qr_ws_bytes<tnh::c128>(m, n, mpl);
  }
  return 0;
}

extern "C" int tnh_qr(int dtype, int batch, int m, int n, const void* A, long lda, long sA,
                      void* Q, long ldq, long sQ, void* R, long ldr, long sR, int* info, void* ws,
                      size_t ws_bytes, tnh_stream_t stream) {
  TNH_REQUIRE(dtype == TNH_F64 || dtype == TNH_C128, TNH_ERR_UNSUPPORTED,
              "tnh_qr: dtype %d (float64 or complex128 only)", dtype);
  const int mmax = dtype == TNH_F64 ? tnh::qr_max_m<double>() : tnh::qr_max_m<tnh::c128>();
  TNH_REQUIRE(n >= 1 && n <= m && m <= mmax, TNH_ERR_UNSUPPORTED,
              "tnh_qr: %d x %d outside 1 <= n <= m <= %d", m, n, mmax);
  TNH_REQUIRE(batch >= 1 && A && Q && R && lda >= n && ldq >= n && ldr >= n &&
                  (batch == 1 || (sA >= lda * (m - 1) + n && sQ >= ldq * (m - 1) + n &&
                                  sR >= ldr * (n - 1) + n)),
              TNH_ERR_ARG, "tnh_qr: bad arguments (batch=%d m=%d n=%d lda=%ld ldq=%ld ldr=%ld)",
              batch, m, n, lda, ldq, ldr);
  const size_t need = tnh_qr_ws_bytes(dtype, batch, m, n);
  TNH_REQUIRE(need > 0, TNH_ERR_HIP, "tnh_qr: no workspace size on this device");
  TNH_REQUIRE(ws && ws_bytes >= need, TNH_ERR_WS, "tnh_qr: workspace %zu < %zu bytes", ws_bytes,
              need);
  hipStream_t st = static_cast<hipStream_t>(stream);
  int mpl = 1;
  if (dtype == TNH_F64) {
    int rc = tnh::qr_mpl<double>(m, n, batch, &mpl);
    if (rc != TNH_OK) return rc;
    return tnh::qr_run<double>(dtype, batch, m, n, mpl, static_cast<const double*>(A), lda, sA,
                               static_cast<double*>(Q), ldq, sQ, static_cast<double*>(R), ldr, sR,
                               info, ws, st);
  }
  int rc = tnh::qr_mpl<tnh::c128>(m, n, batch, &mpl);
  if (rc != TNH_OK) return rc;
  return tnh::qr_run<tnh::c128>(dtype, batch, m, n, mpl, static_cast<const tnh::c128*>(A), lda,
                                sA, static_cast<tnh::c128*>(Q), ldq, sQ,
                                static_cast<tnh::c128*>(R), ldr, sR, info, ws, st);
}
