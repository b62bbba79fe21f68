// The two-site effective Hamiltonian and the environment updates as K-major MFMA GEMMs joined
// by small "mix" kernels that apply the MPO tensors.
//
// Two-site matvec (tnkit/_dmrg.py:172-189, apply_two_site_fused), bonds a -> b, d1, d2, wl, wr:
//   GEMM1  X1[(i j b)][(p w)]  = sum_q theta[q][(i j b)] * LT[q][(p w)]      LT[q][p][w] = L[p][w][q]
//   S-mix  X2[r][(b p)][(s t)] = sum_{w i j} S[(s t r)][(w i j)] * X1[(i j)][(b p)][w]
//   GEMM3  out[(p s t)][d]     = sum_{(r b)} X2[(r b)][(p s t)] * RT[(r b)][d]  RT[r][b][d] = R[d][r][b]
// Left environment (tnkit/_dmrg.py:216-222), L(a, wl, a), A(a, d, c), W(wl, wr, d, d):
//   GEMM   T1[(w q)][(s e)]   = sum_p L[p][(w q)] * A[p][(s e)]
//   W-mix  T2[(q y)][(e r)]   = sum_{w s} W[w][r][s][y] * T1[w][q][s][e]
//   GEMM   L'[(e r)][f]       = sum_{(q y)} T2[(q y)][(e r)] * A[(q y)][f]
// Right environment (tnkit/_dmrg.py:225-231), R(b, wr, b), B(c, d, b):
//   GEMM   T1[(r y)][(u s)]   = sum_x R[x][(r y)] * Bx[x][(u s)]               Bx[x][u][s] = B[u][s][x]
//   W-mix  T2[(p y)][(u w)]   = sum_{r s} W[w][r][s][p] * T1[r][y][u][s]
//   GEMM   R'[(u w)][v]       = sum_{(p y)} T2[(p y)][(u w)] * Bp[(p y)][v]     Bp[p][y][v] = B[v][p][y]
// Real (float64, float32) and complex (complex128, complex64) elements; conj(A) and conj(B) are
// taken inside the first GEMM of each environment update. Every entry has a batched form: k
// independent problems of one shape with element strides between them (the legs of the batched
// time evolution), one launch per kernel for all k.
#include <type_traits>

#include "gemm_tn.h"

namespace tnh {

// ---------------------------------------------------------------------------------------------
// small kernels
// ---------------------------------------------------------------------------------------------

// out[j0][j1][j2] = in[j0*s0 + j1*s1 + j2*s2]   (a 3-axis permutation given by input strides);
// batch element z = blockIdx.z at element offsets z*bin, z*bout
template <typename T>
__device__ __forceinline__ T cj(T v) {
  return v;
}
template <typename T>
__device__ __forceinline__ cpx<T> cj(cpx<T> v) {
  return conj(v);
}

template <typename T, bool CJ = false>
__global__ void permute3_kernel(long n0, long n1, long n2, long s0, long s1, long s2,
                                const T* __restrict__ in, long bin, T* __restrict__ out, long bout) {
  const long total = n0 * n1 * n2;
  in += blockIdx.z * bin;
  out += blockIdx.z * bout;
  for (long o = blockIdx.x * (long)blockDim.x + threadIdx.x; o < total;
       o += (long)gridDim.x * blockDim.x) {
    const long j2 = o % n2;
    const long r = o / n2;
    const long j1 = r % n1;
    const long j0 = r / n1;
    const T v = in[j0 * s0 + j1 * s1 + j2 * s2];
    if constexpr (CJ) {
      out[o] = cj(v);
    } else {
      out[o] = v;
    }
  }
}

// S of the adjoint operator: S'[(s t r)][(w i j)] = conj(sum_m W1[w][m][i][s] W2[m][r][j][t])
template <typename T>
__global__ void build_s_adjoint_kernel(int wl, int wm, int wr, int d1, int d2,
                                       const T* __restrict__ W1, const T* __restrict__ W2,
                                       T* __restrict__ S) {
  const int ncol = wl * d1 * d2;
  const int total = d1 * d2 * wr * ncol;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < total; o += gridDim.x * blockDim.x) {
    const int col = o % ncol, row = o / ncol;
    const int j = col % d2, i = (col / d2) % d1, w = col / (d1 * d2);
    const int r = row % wr, t = (row / wr) % d2, s = row / (wr * d2);
    T acc = T(0);
    for (int m = 0; m < wm; ++m)
      acc += W1[((w * wm + m) * d1 + i) * d1 + s] * W2[((m * wr + r) * d2 + j) * d2 + t];
    S[o] = cj(acc);
  }
}

// x = conj(x) in place (complex element types)
template <typename T>
__global__ void conj_kernel(long n, T* __restrict__ x) {
  for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
    x[i] = cj(x[i]);
}

// S[(s t r)][(w i j)] = sum_m W1[w][m][s][i] * W2[m][r][t][j]   (batch element blockIdx.z)
template <typename T>
__global__ void build_s_kernel(int wl, int wm, int wr, int d1, int d2, const T* __restrict__ W1,
                               long b1, const T* __restrict__ W2, long b2, T* __restrict__ S,
                               long bs) {
  W1 += blockIdx.z * b1;
  W2 += blockIdx.z * b2;
  S += blockIdx.z * bs;
  const int ncol = wl * d1 * d2;
  const int total = d1 * d2 * wr * ncol;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < total; o += gridDim.x * blockDim.x) {
    const int col = o % ncol, row = o / ncol;
    const int j = col % d2, i = (col / d2) % d1, w = col / (d1 * d2);
    const int r = row % wr, t = (row / wr) % d2, s = row / (wr * d2);
    T acc = T(0);
    for (int m = 0; m < wm; ++m)
      acc += W1[((w * wm + m) * d1 + s) * d1 + i] * W2[((m * wr + r) * d2 + t) * d2 + j];
    S[o] = acc;
  }
}

// Mx[(o1 o2)][(i1 i2)] = W[w][r][s][y] with the four MPO axes routed to (o1, o2, i1, i2) by
// `which`: 0 (left env)  Mx[(y r)][(w s)];   1 (right env)  Mx[(p w)][(r s)] = W[w][r][s][p].
template <typename T>
__global__ void build_wmix_kernel(int which, int wl, int wr, int d, const T* __restrict__ W,
                                  long bw, T* __restrict__ Mx, long bm) {
  W += blockIdx.z * bw;
  Mx += blockIdx.z * bm;
  const int total = wl * wr * d * d;
  for (int o = blockIdx.x * blockDim.x + threadIdx.x; o < total; o += gridDim.x * blockDim.x) {
    int w, r, s, y;
    if (which == 0) {  // o = ((y*wr + r)*wl + w)*d + s
      s = o % d;
      w = (o / d) % wl;
      r = (o / (d * wl)) % wr;
      y = o / (d * wl * wr);
    } else {  // o = ((p*wl + w)*wr + r)*d + s, p -> y
      s = o % d;
      r = (o / d) % wr;
      w = (o / (d * wr)) % wl;
      y = o / (d * wr * wl);
    }
    Mx[o] = W[((w * wr + r) * d + s) * d + y];
  }
}

// Generic small-matrix mix over a 2-level group index (g1, g2):
//   out[g1*os1 + g2*os2 + jo1*ojs1 + jo2*ojs2] = sum_ji Mx[jo][ji] * in[g1*is1 + g2*is2 + ji1*ijs1 + ji2*ijs2]
// with jo = jo1*njo2 + jo2, ji = ji1*nji2 + ji2. One thread per group and chunk of CO outputs: the
// thread loads its group's inputs once (CI at a time) into registers and writes its outputs;
// g2 runs fastest, so a wave reads and writes contiguous runs. Mx sits in LDS, and its zero
// entries (most of an MPO's) are skipped by a branch that is uniform across the workgroup.
struct MixArgs {
  long G1, G2;
  int njo, njo2, nji, nji2;
  long is1, is2, ijs1, ijs2;
  long os1, os2, ojs1, ojs2;
  long bm, bin, bout;  // batch strides of Mx, in and out (batch element blockIdx.z)
};

// outputs / inputs per register chunk: 32 floats or 16 doubles stay in registers without spills
template <typename T>
struct MixChunk {
  static constexpr int value = 32;
};
template <>
struct MixChunk<double> {
  static constexpr int value = 16;
};
template <>
struct MixChunk<c64> {
  static constexpr int value = 16;
};
template <>
struct MixChunk<c128> {
  static constexpr int value = 8;
};
constexpr int kMixLds = 2048;  // Mx entries held in LDS (larger operators read global memory)

template <typename TC, typename TI, typename TO>
__global__ void __launch_bounds__(256) mix_kernel(MixArgs a, const TC* __restrict__ Mx,
                                                  const TI* __restrict__ in, TO* __restrict__ out) {
  constexpr int kMixChunk = MixChunk<TC>::value;
  Mx += blockIdx.z * a.bm;
  in += blockIdx.z * a.bin;
  out += blockIdx.z * a.bout;
  __shared__ __align__(16) unsigned char ms_raw[kMixLds * sizeof(TC)];
  TC* ms = reinterpret_cast<TC*>(ms_raw);
  const int nm = a.njo * a.nji;
  const bool in_lds = nm <= kMixLds;
  if (in_lds)
    for (int k = threadIdx.x; k < nm; k += blockDim.x) ms[k] = Mx[k];
  __syncthreads();
  const TC* M = in_lds ? ms : Mx;
  const long G = a.G1 * a.G2;
  const int nco = (a.njo + kMixChunk - 1) / kMixChunk;
  const long total = G * nco;
  for (long tid = blockIdx.x * (long)blockDim.x + threadIdx.x; tid < total;
       tid += (long)gridDim.x * blockDim.x) {
    const long g = tid % G;
    const int co = static_cast<int>(tid / G);
    const long g2 = g % a.G2;
    const long g1 = g / a.G2;
    const long ibase = g1 * a.is1 + g2 * a.is2;
    const int jo0 = co * kMixChunk;
    TC acc[kMixChunk];
#pragma unroll
    for (int o = 0; o < kMixChunk; ++o) acc[o] = TC(0);
    for (int ji0 = 0; ji0 < a.nji; ji0 += kMixChunk) {
      TC x[kMixChunk];
#pragma unroll
      for (int i = 0; i < kMixChunk; ++i) {
        const int ji = ji0 + i;
        if (ji < a.nji) {
          const int ji1 = ji / a.nji2, ji2 = ji - ji1 * a.nji2;
          x[i] = static_cast<TC>(in[ibase + ji1 * a.ijs1 + ji2 * a.ijs2]);
        } else {
          x[i] = TC(0);
        }
      }
#pragma unroll
      for (int o = 0; o < kMixChunk; ++o) {
        const int jo = jo0 + o;
        if (jo < a.njo) {
          const TC* mrow = M + (long)jo * a.nji + ji0;
#pragma unroll
          for (int i = 0; i < kMixChunk; ++i) {
            if (ji0 + i < a.nji) {
              const TC m = mrow[i];
              if (m != TC(0)) acc[o] += m * x[i];
            }
          }
        }
      }
    }
    const long obase = g1 * a.os1 + g2 * a.os2;
#pragma unroll
    for (int o = 0; o < kMixChunk; ++o) {
      const int jo = jo0 + o;
      if (jo < a.njo) {
        const int jo1 = jo / a.njo2, jo2 = jo - jo1 * a.njo2;
        out[obase + jo1 * a.ojs1 + jo2 * a.ojs2] = static_cast<TO>(acc[o]);
      }
    }
  }
}

// LDS-staged form of mix_kernel for operators with nji, njo <= kStageMax. A workgroup takes kStageG
// consecutive g2 at one g1: it copies the inputs of those groups into LDS in the order in which
// they lie in memory (mode 0: g2 fastest; 1: ji1 fastest, then g2; 2: ji2 fastest, then g2), each
// thread computes the outputs of one group from LDS, and the outputs leave through LDS in memory
// order the same way. Every global access of a wave is then a contiguous run.
constexpr int kStageG = 256;
// LDS row stride of the staging buffer: one more than kStageG, so that lanes holding the same
// group t in different rows (the strided load and store orders) fall in different banks
constexpr int kStageLd = kStageG + 1;
constexpr int kStageMax = 32;
template <typename T>
struct StageMax {  // register budget of the staged kernel: 32 real or 16 complex per thread
  static constexpr int value = real_of<T>::cplx ? 16 : kStageMax;
};

// Unsigned division by a divisor fixed for the whole launch: q / d = umulhi(q, ceil(2^32 / d)),
// exact whenever q * d < 2^32 (the rounding excess of the multiplier, below d, then stays below
// one unit in the last place). The staging decode divides indices below kStageMax * kStageG by
// divisors up to kStageMax, far inside that range; a general integer division costs tens of VALU
// instructions per element, which made the staging loops instruction-bound.
struct FastDiv {
  unsigned d, m;
  __device__ explicit FastDiv(unsigned d_)
      : d(d_), m(d_ > 1 ? static_cast<unsigned>((0x100000000ull + d_ - 1) / d_) : 0u) {}
  __device__ __forceinline__ unsigned div(unsigned q) const { return d > 1 ? __umulhi(q, m) : q; }
};

// (j, t) of position L of the staging order: j = j1 * n2 + j2 is the in-group index (j1 < n1,
// j2 < n2), t the group within the workgroup.
__device__ __forceinline__ void stage_decode(int mode, unsigned L, const FastDiv& n1,
                                             const FastDiv& n2, int& j, int& t) {
  if (mode == 1) {  // L = (j2 * G + t) * n1 + j1
    const unsigned r = n1.div(L);
    const int j1 = static_cast<int>(L - r * n1.d);
    t = static_cast<int>(r % kStageG);
    j = j1 * static_cast<int>(n2.d) + static_cast<int>(r / kStageG);
  } else if (mode == 2) {  // L = (j1 * G + t) * n2 + j2
    const unsigned r = n2.div(L);
    const int j2 = static_cast<int>(L - r * n2.d);
    t = static_cast<int>(r % kStageG);
    j = static_cast<int>(r / kStageG) * static_cast<int>(n2.d) + j2;
  } else {  // L = j * G + t
    j = static_cast<int>(L / kStageG);
    t = static_cast<int>(L % kStageG);
  }
}

// Copies the inputs of the kStageG groups g20.. at one g1 into LDS rows buf[ji][t] in the order
// in which they lie in memory. All of a thread's loads are issued before any LDS store, so up to
// kMax are in flight. Offsets within the workgroup's window fit 32 bits (checked on the host).
template <int kMax, typename TC, typename TI>
__device__ __forceinline__ void stage_in(const MixArgs& a, int mode, const TI* __restrict__ in,
                                         long ib, long g20, TC* buf) {
  const FastDiv n2(static_cast<unsigned>(a.nji2));
  const FastDiv n1(static_cast<unsigned>(a.nji / a.nji2));
  const TI* base = in + ib + g20 * a.is2;
  const int is2 = static_cast<int>(a.is2), s1 = static_cast<int>(a.ijs1),
            s2 = static_cast<int>(a.ijs2);
  const int tmax = static_cast<int>(a.G2 - g20 < kStageG ? a.G2 - g20 : kStageG);
  TC v[kMax];
  int dst[kMax];
#pragma unroll
  for (int i = 0; i < kMax; ++i) {
    v[i] = TC(0);
    dst[i] = -1;
    if (i < a.nji) {
      int ji, t;
      stage_decode(mode, threadIdx.x + i * kStageG, n1, n2, ji, t);
      dst[i] = ji * kStageLd + t;
      if (t < tmax) {
        const int ji1 = static_cast<int>(n2.div(ji)), ji2 = ji - ji1 * a.nji2;
        v[i] = static_cast<TC>(base[t * is2 + ji1 * s1 + ji2 * s2]);
      }
    }
  }
#pragma unroll
  for (int i = 0; i < kMax; ++i)
    if (dst[i] >= 0) buf[dst[i]] = v[i];
}

// Writes the LDS rows buf[jo][t] of the kStageG groups g20.. at one g1 out in memory order.
template <typename TC, typename TO>
__device__ __forceinline__ void stage_out(const MixArgs& a, int mode, const TC* buf,
                                          TO* __restrict__ out, long ob, long g20) {
  const FastDiv n2(static_cast<unsigned>(a.njo2));
  const FastDiv n1(static_cast<unsigned>(a.njo / a.njo2));
  TO* base = out + ob + g20 * a.os2;
  const int os2 = static_cast<int>(a.os2), s1 = static_cast<int>(a.ojs1),
            s2 = static_cast<int>(a.ojs2);
  const int tmax = static_cast<int>(a.G2 - g20 < kStageG ? a.G2 - g20 : kStageG);
  const int nout = a.njo * kStageG;
  for (int L = threadIdx.x; L < nout; L += blockDim.x) {
    int jo, tt;
    stage_decode(mode, L, n1, n2, jo, tt);
    if (tt < tmax) {
      const int jo1 = static_cast<int>(n2.div(jo)), jo2 = jo - jo1 * a.njo2;
      base[tt * os2 + jo1 * s1 + jo2 * s2] = static_cast<TO>(buf[jo * kStageLd + tt]);
    }
  }
}

template <typename TC, typename TI, typename TO>
__global__ void __launch_bounds__(256) mix_staged_kernel(MixArgs a, int in_mode, int out_mode,
                                                         const TC* __restrict__ Mx,
                                                         const TI* __restrict__ in,
                                                         TO* __restrict__ out) {
  constexpr int kMax = StageMax<TC>::value;
  Mx += blockIdx.z * a.bm;
  in += blockIdx.z * a.bin;
  out += blockIdx.z * a.bout;
  // dynamic LDS: the operator (njo x nji), then max(nji, njo) rows of kStageG values
  extern __shared__ __align__(16) unsigned char stage_lds[];
  TC* ms = reinterpret_cast<TC*>(stage_lds);
  const int nm = a.njo * a.nji;
  TC* buf = ms + ((nm + 1) / 2) * 2;
  for (int k = threadIdx.x; k < nm; k += blockDim.x) ms[k] = Mx[k];
  const long g1 = blockIdx.y;
  const long g20 = (long)blockIdx.x * kStageG;
  stage_in<kMax>(a, in_mode, in, g1 * a.is1, g20, buf);
  __syncthreads();
  TC acc[kMax];
#pragma unroll
  for (int o = 0; o < kMax; ++o) acc[o] = TC(0);
  const int t = threadIdx.x;
  for (int ji = 0; ji < a.nji; ++ji) {
    const TC x = buf[ji * kStageLd + t];
#pragma unroll
    for (int o = 0; o < kMax; ++o) {
      if (o < a.njo) {
        const TC m = ms[o * a.nji + ji];
        if (m != TC(0)) acc[o] += m * x;
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int o = 0; o < kMax; ++o)
    if (o < a.njo) buf[o * kStageLd + t] = acc[o];
  __syncthreads();
  stage_out(a, out_mode, buf, out, g1 * a.os1, g20);
}

// The staged mix on the matrix cores (real types). The mix of the kStageG groups of a workgroup
// is the product Out^T (groups x njo) = In^T (groups x nji) * Mx^T (nji x njo): each wave takes
// 64 groups as four 16-row MFMA tiles, K = nji in steps of 4 (at most 8), N = njo in one or two
// 16-column tiles. The Mx^T fragments are loaded once per lane from global memory (the operator
// is a few hundred entries and stays in cache); In^T fragments are read from the LDS rows that
// stage_in filled, and the results go back to LDS rows for stage_out. A scalar FMA form spends
// several VALU instructions per operator entry and group; here a wave issues nks * nns MFMAs per
// 16 groups, so the kernel is bound by its HBM streams.
template <typename TC, typename TI, typename TO>
__global__ void __launch_bounds__(256) mix_mfma_kernel(MixArgs a, int in_mode, int out_mode,
                                                       const TC* __restrict__ Mx,
                                                       const TI* __restrict__ in,
                                                       TO* __restrict__ out) {
  using MF = Mfma<TC>;
  using acc_t = typename MF::acc_t;
  Mx += blockIdx.z * a.bm;
  in += blockIdx.z * a.bin;
  out += blockIdx.z * a.bout;
  extern __shared__ __align__(16) unsigned char stage_lds[];
  TC* buf = reinterpret_cast<TC*>(stage_lds);
  const long g1 = blockIdx.y;
  const long g20 = (long)blockIdx.x * kStageG;
  stage_in<kStageMax>(a, in_mode, in, g1 * a.is1, g20, buf);
  const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6, l16 = lane & 15, kq = lane >> 4;
  const int nji = a.nji, njo = a.njo;
  const int nks = (nji + 3) >> 2, nns = (njo + 15) >> 4;
  TC bfr[8][2];
#pragma unroll
  for (int ks = 0; ks < 8; ++ks)
#pragma unroll
    for (int ns = 0; ns < 2; ++ns) {
      const int k = 4 * ks + kq, jo = 16 * ns + l16;
      bfr[ks][ns] = (k < nji && jo < njo) ? Mx[jo * nji + k] : TC(0);
    }
  __syncthreads();
  acc_t acc[4][2];
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int ns = 0; ns < 2; ++ns) acc[mt][ns] = acc_t{};
  const TC* col = buf + 64 * wave + l16;
#pragma unroll
  for (int ks = 0; ks < 8; ++ks) {
    if (ks < nks) {
      const int k = 4 * ks + kq;
#pragma unroll
      for (int mt = 0; mt < 4; ++mt) {
        const TC x = k < nji ? col[k * kStageLd + 16 * mt] : TC(0);
        acc[mt][0] = MF::run(x, bfr[ks][0], acc[mt][0]);
        if (nns > 1) acc[mt][1] = MF::run(x, bfr[ks][1], acc[mt][1]);
      }
    }
  }
  __syncthreads();
#pragma unroll
  for (int mt = 0; mt < 4; ++mt)
#pragma unroll
    for (int ns = 0; ns < 2; ++ns) {
      const int jo = 16 * ns + l16;
      if (ns < nns && jo < njo) {
#pragma unroll
        for (int r = 0; r < 4; ++r)
          buf[jo * kStageLd + 64 * wave + 16 * mt + MF::row(kq, r)] = acc[mt][ns][r];
      }
    }
  __syncthreads();
  stage_out(a, out_mode, buf, out, g1 * a.os1, g20);
}

namespace {

inline unsigned grid_for(long total, int threads = 256, long cap = 1L << 16) {
  long g = (total + threads - 1) / threads;
  if (g < 1) g = 1;
  if (g > cap) g = cap;
  return static_cast<unsigned>(g);
}

template <typename T, bool CJ = false>
int permute3(int batch, long n0, long n1, long n2, long s0, long s1, long s2, const void* in,
             long bin, void* out, long bout, hipStream_t st) {
  const long total = n0 * n1 * n2;
  if (total == 0 || batch == 0) return TNH_OK;
  hipLaunchKernelGGL((permute3_kernel<T, CJ>), dim3(grid_for(total), 1, batch), dim3(256), 0, st, n0, n1,
                     n2, s0, s1, s2, static_cast<const T*>(in), bin, static_cast<T*>(out), bout);
  TNH_CHECK_LAUNCH("permute3");
  return TNH_OK;
}

// memory order of a 2-level index block: 0 = the group index g2 is the unit stride, 1 = the
// first component is, with g2 stepping over all of it, 2 = the same for the second component;
// -1 = none of these
inline int stage_mode(long s_g2, long s1, long s2, int n, int n2) {
  const int n1 = n / n2;
  if (s_g2 == 1) return 0;
  if (s1 == 1 && s_g2 == n1) return 1;
  if (s2 == 1 && s_g2 == n2) return 2;
  return -1;
}

template <typename TC, typename TI, typename TO>
int mix(const MixArgs& a, int batch, const void* Mx, const void* in, void* out, hipStream_t st) {
  if (a.G1 * a.G2 == 0 || batch == 0) return TNH_OK;
  const int im = stage_mode(a.is2, a.ijs1, a.ijs2, a.nji, a.nji2);
  const int om = stage_mode(a.os2, a.ojs1, a.ojs2, a.njo, a.njo2);
  const int rows = a.nji > a.njo ? a.nji : a.njo;
  const size_t lds = (((size_t)a.njo * a.nji + 1) / 2 * 2 + (size_t)rows * kStageLd) * sizeof(TC);
  constexpr int kMax = StageMax<TC>::value;
  // the staged kernels index within one workgroup's window in 32 bits
  const long win_in = (long)(kStageG - 1) * a.is2 + (long)(a.nji / a.nji2 - 1) * a.ijs1 +
                      (long)(a.nji2 - 1) * a.ijs2;
  const long win_out = (long)(kStageG - 1) * a.os2 + (long)(a.njo / a.njo2 - 1) * a.ojs1 +
                       (long)(a.njo2 - 1) * a.ojs2;
  const bool win_ok = win_in < (1L << 31) && win_out < (1L << 31) && a.is2 >= 0 &&
                      a.ijs1 >= 0 && a.ijs2 >= 0 && a.os2 >= 0 && a.ojs1 >= 0 && a.ojs2 >= 0;
  if (a.nji <= kMax && a.njo <= kMax && im >= 0 && om >= 0 && a.G1 <= 65535 && lds <= 65536 &&
      win_ok) {
    dim3 grid(static_cast<unsigned>((a.G2 + kStageG - 1) / kStageG), static_cast<unsigned>(a.G1),
              batch);
    if constexpr (!real_of<TC>::cplx) {
      const size_t lds_mfma = (size_t)rows * kStageLd * sizeof(TC);
      hipLaunchKernelGGL((mix_mfma_kernel<TC, TI, TO>), grid, dim3(kStageG), lds_mfma, st, a, im,
                         om, static_cast<const TC*>(Mx), static_cast<const TI*>(in),
                         static_cast<TO*>(out));
      TNH_CHECK_LAUNCH("mix_mfma");
      return TNH_OK;
    }
    hipLaunchKernelGGL((mix_staged_kernel<TC, TI, TO>), grid, dim3(kStageG), lds, st, a, im, om,
                       static_cast<const TC*>(Mx), static_cast<const TI*>(in),
                       static_cast<TO*>(out));
    TNH_CHECK_LAUNCH("mix_staged");
    return TNH_OK;
  }
  constexpr int kMixChunk = MixChunk<TC>::value;
  const long total = a.G1 * a.G2 * ((a.njo + kMixChunk - 1) / kMixChunk);
  hipLaunchKernelGGL((mix_kernel<TC, TI, TO>), dim3(grid_for(total, 256, 1L << 16), 1, batch),
                     dim3(256), 0, st, a, static_cast<const TC*>(Mx), static_cast<const TI*>(in),
                     static_cast<TO*>(out));
  TNH_CHECK_LAUNCH("mix");
  return TNH_OK;
}

size_t dsize(int dtype) {
  switch (dtype) {
    case TNH_F64: return 8;
    case TNH_F32: return 4;
    case TNH_C128: return 16;
    case TNH_C64: return 8;
    default: return 0;
  }
}

#define TNH_TRY(x)                 \
  do {                             \
    int rc_ = (x);                 \
    if (rc_ != TNH_OK) return rc_; \
  } while (0)

// ---- two-site ----------------------------------------------------------------------------------

struct Strides4 {  // element strides between batch elements
  long a, b, c, d;
};

template <typename T>
int fused_two_site(int batch, int a, int wl, int wm, int wr, int d1, int d2, int b, const void* L,
                   long sL, const void* W1, long sW1, const void* W2, long sW2, const void* R,
                   long sR, void* LT, long sLT, void* S, long sS, void* RT, long sRT,
                   hipStream_t st) {
  // LT[q][p][w] = L[p][w][q]: L strides (wl*a, a, 1) for (p, w, q)
  TNH_TRY(permute3<T>(batch, a, a, wl, 1, (long)wl * a, a, L, sL, LT, sLT, st));
  // RT[r][b][d] = R[d][r][b]: R strides (wr*b, b, 1) for (d, r, b)
  TNH_TRY(permute3<T>(batch, wr, b, b, b, 1, (long)wr * b, R, sR, RT, sRT, st));
  const int total = d1 * d2 * wr * wl * d1 * d2;
  hipLaunchKernelGGL(build_s_kernel<T>, dim3(grid_for(total), 1, batch), dim3(256), 0, st, wl, wm,
                     wr, d1, d2, static_cast<const T*>(W1), sW1, static_cast<const T*>(W2), sW2,
                     static_cast<T*>(S), sS);
  TNH_CHECK_LAUNCH("build_s");
  return TNH_OK;
}

// stages: bit 0 GEMM1, bit 1 S-mix, bit 2 GEMM3 (7 = the whole matvec). A partial run reads and
// writes the workspace as the whole one does; it exists to time one stage at a time.
// theta (float64) -> float32 ahead of the mixed matvec's GEMM1: the GEMM then stages 4-byte
// operands (its tile loads and LDS stores are those of the all-float32 kernel), for one pass
// over theta (12 bytes per element).
__global__ void __launch_bounds__(256) narrow_kernel(long n, const double* __restrict__ x,
                                                     float* __restrict__ y) {
  const long stride = (long)gridDim.x * 256 * 2;
  for (long i = ((long)blockIdx.x * 256 + threadIdx.x) * 2; i < n; i += stride) {
    if (i + 1 < n) {
      const double2 v = *reinterpret_cast<const double2*>(x + i);
      *reinterpret_cast<float2*>(y + i) = make_float2(static_cast<float>(v.x),
                                                      static_cast<float>(v.y));
    } else {
      y[i] = static_cast<float>(x[i]);
    }
  }
}


// ---- GEMM1 with the MPO mix in its epilogue (float32) -----------------------------------------
// For d1 = d2 = 2 the first two stages of the matvec fuse: a GEMM1 output tile of 128 rows
// (the four (i, j) blocks of 32 consecutive b) by 32 WL columns (32 consecutive p, all WL
// channels w) holds, for each of its 32 x 32 (b, p) pairs, the 4 WL inputs of the MPO mix, so
// the tile is mixed in the epilogue and written straight to X2[r][(b p)][(s t)]; X1 never goes
// to HBM (it was 84 MB written and read again in float32 at chi = 1024). The main loop is the
// staged K-major loop of gemm_tn_kernel with the A rows gathered from the four (i, j)
// segments. The epilogue runs in two halves of 16 p (the two wave columns), through an LDS tile
// that aliases the main loop's buffers: acc -> LDS, then each thread mixes two (b, p) pairs with
// the 4 WR x 4 WL operator S and stores 4 (s, t) values per channel r.
struct Gemm1MixArgs {
  int K, B, P, WR;
  const void* A;   // theta in the compute type: K rows of 4 B (i j b)
  long lda, sA;
  const void* Bm;  // LT: K rows of P WL (p w)
  long ldb, sB;
  const void* S;   // (s t r) x (w i j), row-major, 4 WR x 4 WL
  long sS;
  void* X2;        // (r, b, p, s t)
  long sX2;
};

template <typename T, int WL, int NWM>
__global__ void __launch_bounds__(NWM * 128, NWM == 2 ? 3 : 1)
    gemm1_mix_kernel(Gemm1MixArgs g) {
  constexpr int NT = NWM * 128;                     // NWM x 2 waves
  using MF = Mfma<T>;
  using acc_t = typename MF::acc_t;
  constexpr int BM = 128, BN = 32 * WL, BK = 16, PAD = 16;
  constexpr int LDA_S = BM + PAD, LDB_S = BN + PAD;
  constexpr int FM = 128 / NWM / 16, FN = WL;       // waves of (128 / NWM) x 16 WL
  constexpr int VA = 16 / sizeof(T);                // A: 16-byte vectors
  constexpr int VB = 8 / sizeof(T);                 // B: 8-byte vectors (BN is a multiple of 32)
  constexpr int NA = BK * BM / (VA * NT), NB = BK * BN / (VB * NT);
  static_assert(NA * VA * NT == BK * BM && NB * VB * NT == BK * BN, "gemm1_mix tile loads");
  constexpr int ELD = 16 * WL + 1;                  // epilogue tile row (one half), odd stride
  constexpr int MAIN = 2 * BK * (LDA_S + LDB_S);
  constexpr int EPI = BM * ELD;
  __shared__ __align__(16) T lds[MAIN > EPI ? MAIN : EPI];
  __shared__ T Ss[4 * 8 * 4 * WL];
  T* As0 = lds;
  T* Bs0 = lds + 2 * BK * LDA_S;

  const int B = g.B, P = g.P, K = g.K, WR = g.WR;
  const int nbm = B / 32, nbn = P / 32;
  int tm, tn;
  tile_of(blockIdx.x, nbm, nbn, tm, tn);
  const int b0 = tm * 32, p0 = tn * 32;
  const long bz = blockIdx.y;
  const T* __restrict__ A = static_cast<const T*>(g.A) + bz * g.sA;
  const T* __restrict__ Bm = static_cast<const T*>(g.Bm) + bz * g.sB;
  const T* __restrict__ Sg = static_cast<const T*>(g.S) + bz * g.sS;
  T* __restrict__ X2 = static_cast<T*>(g.X2) + bz * g.sX2;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1, row16 = lane & 15, kq = lane >> 4;
  for (int e = t; e < 4 * WR * 4 * WL; e += NT) Ss[e] = Sg[e];

  typedef T va_t __attribute__((ext_vector_type(VA)));
  typedef T vb_t __attribute__((ext_vector_type(VB)));
  va_t ra[NA];
  vb_t rb[NB];
  auto load = [&](int k0) {
#pragma unroll
    for (int v = 0; v < NA; ++v) {
      const int e = t + v * NT, kr = e / (BM / VA), mc = (e % (BM / VA)) * VA;
      const int k = k0 + kr;
      // columns mc .. mc + VA - 1 of the tile: segment ij = mc / 32, b = b0 + mc % 32
      if (k < K)
        ra[v] = *reinterpret_cast<const va_t*>(A + (long)k * g.lda + (long)(mc >> 5) * B + b0 +
                                              (mc & 31));
      else
        ra[v] = va_t{};
    }
#pragma unroll
    for (int v = 0; v < NB; ++v) {
      const int e = t + v * NT, kr = e / (BN / VB), nc = (e % (BN / VB)) * VB;
      const int k = k0 + kr;
      if (k < K)
        rb[v] = *reinterpret_cast<const vb_t*>(Bm + (long)k * g.ldb + (long)p0 * WL + nc);
      else
        rb[v] = vb_t{};
    }
  };
  auto store = [&](int buf) {
    T* as = As0 + buf * BK * LDA_S;
    T* bs = Bs0 + buf * BK * LDB_S;
#pragma unroll
    for (int v = 0; v < NA; ++v) {
      const int e = t + v * NT, kr = e / (BM / VA), mc = (e % (BM / VA)) * VA;
      *reinterpret_cast<va_t*>(as + kr * LDA_S + mc) = ra[v];
    }
#pragma unroll
    for (int v = 0; v < NB; ++v) {
      const int e = t + v * NT, kr = e / (BN / VB), nc = (e % (BN / VB)) * VB;
      *reinterpret_cast<vb_t*>(bs + kr * LDB_S + nc) = rb[v];
    }
  };

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};
  const int nk = (K + BK - 1) / BK;
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
    const T* as = As0 + cur * BK * LDA_S;
    const T* bs = Bs0 + cur * BK * LDB_S;
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = as[(kk + kq) * LDA_S + wm * (BM / NWM) + i * 16 + row16];
#pragma unroll
      for (int j = 0; j < FN; ++j)
        bf[j] = bs[(kk + kq) * LDB_S + wn * 16 * WL + j * 16 + row16];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MF::run(af[i], bf[j], acc[i][j]);
    }
    if (kt + 1 < nk) store(cur ^ 1);
    __syncthreads();
  }

  // epilogue: two halves of 16 p (wave column h), each through the LDS tile E[128][16 WL]
  T* E = lds;
  const long plane = (long)B * P * 4;  // X2 stride between channels r
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (wn == h) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            E[(wm * (BM / NWM) + i * 16 + MF::row(kq, r)) * ELD + j * 16 + row16] = acc[i][j][r];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 512 / NT; ++q) {
      const int pair = t + q * NT;           // (bb, pp), pp fastest: 32 b x 16 p
      const int bb = pair >> 4, pp = pair & 15;
      T x[WL][4];
#pragma unroll
      for (int w = 0; w < WL; ++w)
#pragma unroll
        for (int ij = 0; ij < 4; ++ij) x[w][ij] = E[(ij * 32 + bb) * ELD + pp * WL + w];
      T* dst = X2 + ((long)(b0 + bb) * P + p0 + 16 * h + pp) * 4;
      for (int r = 0; r < WR; ++r) {
        T y[4];
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const T* srow = Ss + (st * WR + r) * (4 * WL);
          T s = T(0);
#pragma unroll
          for (int w = 0; w < WL; ++w)
#pragma unroll
            for (int ij = 0; ij < 4; ++ij) s += srow[w * 4 + ij] * x[w][ij];
          y[st] = s;
        }
        T* o = dst + r * plane;
        o[0] = y[0];
        o[1] = y[1];
        o[2] = y[2];
        o[3] = y[3];
      }
    }
    __syncthreads();
  }
}

// The same kernel with direct-to-LDS loads: the main loop of gemm_dma_kernel (tiles go from HBM
// into LDS by global_load_lds_dwordx4, no VGPR staging) on the 128 x 32 WL tile of
// gemm1_mix_kernel, BK = 32, two stages. float32 runs 4 waves and two workgroups per CU, so that
// one's epilogue overlaps the other's main loop; float64 runs 8 waves (one workgroup per CU: its
// stages take 144 KB). A rows are rotated by 16 elements per (row mod 4) within their 128
// elements, B rows within their 32 WL (as gemm_dma_kernel does, for distinct LDS banks across
// the four k rows of a fragment read). The epilogue is that of gemm1_mix_kernel.
template <typename T, int WL, int BK>
struct Gemm1MixDmaSmem {
  struct Stages {
    T a0[BK * 128], a1[BK * 128];
    T b0[BK * 32 * WL], b1[BK * 32 * WL];
  };
  union {
    Stages st;
    T E[128 * (16 * WL + 1)];
  };
};

template <typename T, int WL, int NWM>
__global__ void __launch_bounds__(NWM * 128, NWM == 2 ? 2 : 1)
    gemm1_mix_dma_kernel(Gemm1MixArgs g) {
  using MF = Mfma<T>;
  using acc_t = typename MF::acc_t;
  constexpr int NT = NWM * 128;
  constexpr int BM = 128, BN = 32 * WL, BK = 32;
  constexpr int EPC = 16 / sizeof(T), ROT = 16 / EPC;         // elements per chunk; rotation
  constexpr int CPA = BM / EPC, CPB = BN / EPC;               // 16-byte chunks per k row
  constexpr int NJA = BK * CPA / NT, NJB = BK * CPB / NT;     // chunks per thread per tile
  static_assert(NJA * NT == BK * CPA && NJB * NT == BK * CPB, "gemm1_mix_dma tile loads");
  static_assert(3 * ROT <= CPB, "gemm1_mix_dma rotation");
  constexpr int FM = BM / NWM / 16, FN = WL, ELD = 16 * WL + 1;
  __shared__ __align__(16) Gemm1MixDmaSmem<T, WL, BK> sm;
  __shared__ T Ss[4 * 8 * 4 * WL];

  const int B = g.B, P = g.P, K = g.K, WR = g.WR;
  int tm, tn;
  tile_of(blockIdx.x, B / 32, P / 32, tm, tn);
  const int b0 = tm * 32, p0 = tn * 32;
  const long bz = blockIdx.y;
  const T* __restrict__ A = static_cast<const T*>(g.A) + bz * g.sA;
  const T* __restrict__ Bm = static_cast<const T*>(g.Bm) + bz * g.sB;
  const T* __restrict__ Sg = static_cast<const T*>(g.S) + bz * g.sS;
  T* __restrict__ X2 = static_cast<T*>(g.X2) + bz * g.sX2;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave >> 1, wn = wave & 1, row16 = lane & 15, kq = lane >> 4;
  for (int e = t; e < 4 * WR * 4 * WL; e += NT) Ss[e] = Sg[e];

  // element offsets within a K block of this thread's chunks: chunk c lands at LDS position c
  // and holds the chunk of its row rotated back by ROT chunks per (row mod 4)
  long goffA[NJA], goffB[NJB];
#pragma unroll
  for (int j = 0; j < NJA; ++j) {
    const int c = j * NT + t, r = c / CPA, p = c % CPA;
    const int mc = EPC * ((p - ROT * (r & 3)) & (CPA - 1));
    goffA[j] = (long)r * g.lda + (long)(mc >> 5) * B + b0 + (mc & 31);
  }
#pragma unroll
  for (int j = 0; j < NJB; ++j) {
    const int c = j * NT + t, r = c / CPB, p = c % CPB;
    const int nc = EPC * ((p - ROT * (r & 3) + CPB) % CPB);
    goffB[j] = (long)r * g.ldb + (long)p0 * WL + nc;
  }
  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};
  const int nk = K / BK;
  const int aoff = wm * (BM / NWM) + row16 + 16 * kq, boff = wn * 16 * WL + row16 + 16 * kq;
  auto issue = [&](int kt, T* as, T* bs) __attribute__((always_inline)) {
    const T* a = A + (long)kt * BK * g.lda;
    const T* b = Bm + (long)kt * BK * g.ldb;
#pragma unroll
    for (int j = 0; j < NJA; ++j)
      __builtin_amdgcn_global_load_lds(static_cast<const void*>(a + goffA[j]),
                                       (lds_void_t)(as + (j * NT + wave * 64) * EPC), 16, 0, 0);
#pragma unroll
    for (int j = 0; j < NJB; ++j)
      __builtin_amdgcn_global_load_lds(static_cast<const void*>(b + goffB[j]),
                                       (lds_void_t)(bs + (j * NT + wave * 64) * EPC), 16, 0, 0);
  };
  auto compute = [&](const T* as, const T* bs) __attribute__((always_inline)) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = as[(kk + kq) * BM + ((aoff + i * 16) & (BM - 1))];
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int c = boff + j * 16;
        bf[j] = bs[(kk + kq) * BN + (c >= BN ? c - BN : c)];
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MF::run(af[i], bf[j], acc[i][j]);
    }
  };
  // step kt on stage S: tile kt landed (the only loads in flight), barrier (every wave is done
  // with the other stage), prefetch tile kt + 1 into the other stage, compute tile kt
  auto step = [&](auto stage, int kt) __attribute__((always_inline)) {
    constexpr int S = decltype(stage)::value;
    wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    if (kt + 1 < nk) {
      if constexpr (S == 0)
        issue(kt + 1, sm.st.a1, sm.st.b1);
      else
        issue(kt + 1, sm.st.a0, sm.st.b0);
    }
    if constexpr (S == 0)
      compute(sm.st.a0, sm.st.b0);
    else
      compute(sm.st.a1, sm.st.b1);
  };
  if (nk > 0) issue(0, sm.st.a0, sm.st.b0);
  for (int kt = 0; kt < nk; kt += 2) {
    step(std::integral_constant<int, 0>{}, kt);
    if (kt + 1 < nk) step(std::integral_constant<int, 1>{}, kt + 1);
  }
  __syncthreads();

  // epilogue: two halves of 16 p (wave column h), each through the LDS tile E[128][16 WL]
  T* E = sm.E;
  const long plane = (long)B * P * 4;  // X2 stride between channels r
#pragma unroll
  for (int h = 0; h < 2; ++h) {
    if (wn == h) {
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j)
#pragma unroll
          for (int r = 0; r < 4; ++r)
            E[(wm * (BM / NWM) + i * 16 + MF::row(kq, r)) * ELD + j * 16 + row16] = acc[i][j][r];
    }
    __syncthreads();
#pragma unroll
    for (int q = 0; q < 512 / NT; ++q) {
      const int pair = t + q * NT;           // (bb, pp), pp fastest: 32 b x 16 p
      const int bb = pair >> 4, pp = pair & 15;
      T x[WL][4];
#pragma unroll
      for (int w = 0; w < WL; ++w)
#pragma unroll
        for (int ij = 0; ij < 4; ++ij) x[w][ij] = E[(ij * 32 + bb) * ELD + pp * WL + w];
      T* dst = X2 + ((long)(b0 + bb) * P + p0 + 16 * h + pp) * 4;
      for (int r = 0; r < WR; ++r) {
        T y[4];
#pragma unroll
        for (int st = 0; st < 4; ++st) {
          const T* srow = Ss + (st * WR + r) * (4 * WL);
          T s = T(0);
#pragma unroll
          for (int w = 0; w < WL; ++w)
#pragma unroll
            for (int ij = 0; ij < 4; ++ij) s += srow[w * 4 + ij] * x[w][ij];
          y[st] = s;
        }
        T* o = dst + r * plane;
        if constexpr (sizeof(T) == 4) {
          *reinterpret_cast<float4*>(o) = make_float4(y[0], y[1], y[2], y[3]);
        } else {
          *reinterpret_cast<double2*>(o) = make_double2(y[0], y[1]);
          *reinterpret_cast<double2*>(o + 2) = make_double2(y[2], y[3]);
        }
      }
    }
    __syncthreads();
  }
}

template <typename T>
bool try_gemm1_mix(int batch, int a, int d1, int d2, int b, int wl, int wr, const void* theta,
                   long sth, const void* LT, long sLT, const void* S, long sS, void* X2, long sX2,
                   hipStream_t st, int& rc) {
  if (d1 != 2 || d2 != 2 || (wl != 3 && wl != 5) || wr < 1 || wr > 8) return false;
  if (b % 32 || a % 32 || a < 32 || b < 32 || batch < 1 || batch > 65535) return false;
  constexpr int VA = 16 / sizeof(T), VB = 8 / sizeof(T);
  if (reinterpret_cast<uintptr_t>(theta) % 16 || reinterpret_cast<uintptr_t>(LT) % 8 ||
      (batch > 1 && (sth % VA || sLT % VB)))
    return false;
  Gemm1MixArgs g{a, b, a, wr, theta, 4L * b, sth, LT, (long)a * wl, sLT, S, sS, X2, sX2};
  const dim3 grid(static_cast<unsigned>((b / 32) * (a / 32)), static_cast<unsigned>(batch));
  {
    // direct-to-LDS form: 16-byte chunks of LT rows (a wl elements, a % 32 == 0), X2 in 16-byte
    // stores; float32 4 waves of 64 x 16 wl, float64 8 waves of 32 x 16 wl
    constexpr int NWD = sizeof(T) == 4 ? 2 : 4;
    constexpr int EPC = 16 / sizeof(T);
    if (reinterpret_cast<uintptr_t>(LT) % 16 == 0 && reinterpret_cast<uintptr_t>(X2) % 16 == 0 &&
        (batch == 1 || (sth % EPC == 0 && sLT % EPC == 0 && sX2 % EPC == 0))) {
      if (wl == 5)
        hipLaunchKernelGGL((gemm1_mix_dma_kernel<T, 5, NWD>), grid, dim3(NWD * 128), 0, st, g);
      else
        hipLaunchKernelGGL((gemm1_mix_dma_kernel<T, 3, NWD>), grid, dim3(NWD * 128), 0, st, g);
      hipError_t e = hipGetLastError();
      if (e != hipSuccess) {
        set_error("gemm1_mix_dma: launch failed: %s", hipGetErrorString(e));
        rc = TNH_ERR_HIP;
      } else {
        rc = TNH_OK;
      }
      return true;
    }
  }
  // float32: 4 waves of 64 x 16 wl; float64: 8 waves of 32 x 16 wl (half the accumulator
  // registers per lane, so two waves per SIMD fit)
  constexpr int NWM = sizeof(T) == 4 ? 2 : 4;
  if (wl == 5)
    hipLaunchKernelGGL((gemm1_mix_kernel<T, 5, NWM>), grid, dim3(NWM * 128), 0, st, g);
  else
    hipLaunchKernelGGL((gemm1_mix_kernel<T, 3, NWM>), grid, dim3(NWM * 128), 0, st, g);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) {
    set_error("gemm1_mix: launch failed: %s", hipGetErrorString(e));
    rc = TNH_ERR_HIP;
  } else {
    rc = TNH_OK;
  }
  return true;
}

template <typename TC, typename TIO>
int apply_two_site(int batch, int a, int d1, int d2, int b, int wl, int wr, const void* LT,
                   long sLT, const void* S, long sS, const void* RT, long sRT, const void* theta,
                   long sth, void* out, long sout, void* ws, size_t ws_bytes, hipStream_t st,
                   int tile1, int tile3, int stages, const int* live = nullptr) {
  const long P = a, Q = a, B = b, dd = (long)d1 * d2;
  const long nx1 = dd * B * P * wl, nx2 = (long)wr * B * P * dd;
  Arena ar(ws, ws_bytes);
  TC* X1 = ar.take<TC>(nx1 * batch);
  TC* X2 = ar.take<TC>(nx2 * batch);
  // the mixed form narrows theta first when its batches are packed (the workspace query
  // reserves room for it)
  const long nth = Q * dd * B;
  constexpr bool kNarrowR = std::is_same<TC, float>::value && std::is_same<TIO, double>::value;
  constexpr bool kNarrowC = std::is_same<TC, c64>::value && std::is_same<TIO, c128>::value;
  constexpr bool kNarrow = kNarrowR || kNarrowC;
  constexpr long kParts = kNarrowC ? 2 : 1;   // doubles narrowed per element
  TC* Th = nullptr;
  if constexpr (kNarrow) {
    if (batch == 1 || sth == nth) Th = ar.take<TC>(nth * batch);
  }
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "apply_two_site_fused: workspace %zu < %zu bytes", ws_bytes,
              ar.used);
  // GEMM1 and the MPO mix in one kernel when both run here and the shape fits it
  bool fused = false;
  if ((stages & 3) == 3 && tile1 == 0) {
    if constexpr (std::is_same<TC, TIO>::value && !real_of<TC>::cplx) {
      int rc = TNH_OK;
      fused = try_gemm1_mix<TC>(batch, a, d1, d2, b, wl, wr, theta, sth, LT, sLT, S, sS, X2,
                                nx2, st, rc);
      if (fused && rc != TNH_OK) return rc;
    } else if constexpr (kNarrowR) {
      if (Th != nullptr) {
        const long n = nth * batch;
        long gsz = (n / 2 + 255) / 256;
        if (gsz > 4096) gsz = 4096;
        if (gsz < 1) gsz = 1;
        int rc = TNH_OK;
        if (d1 == 2 && d2 == 2 && (wl == 3 || wl == 5) && wr <= 8 && a % 32 == 0 &&
            b % 32 == 0) {
          hipLaunchKernelGGL(narrow_kernel, dim3(static_cast<unsigned>(gsz)), dim3(256), 0, st,
                             n, static_cast<const double*>(theta), Th);
          TNH_CHECK_LAUNCH("narrow");
          fused = try_gemm1_mix<float>(batch, a, d1, d2, b, wl, wr, Th, nth, LT, sLT, S, sS, X2,
                                       nx2, st, rc);
          if (fused && rc != TNH_OK) return rc;
        }
      }
    }
  }
  if ((stages & 1) && !fused) {  // GEMM1: M = d1 d2 b, N = p wl, K = q
    if constexpr (kNarrow) {
      if (Th != nullptr) {
        const long n = nth * batch * kParts;
        long g = (n / 2 + 255) / 256;
        if (g > 4096) g = 4096;
        if (g < 1) g = 1;
        hipLaunchKernelGGL(narrow_kernel, dim3(static_cast<unsigned>(g)), dim3(256), 0, st, n,
                           static_cast<const double*>(theta), reinterpret_cast<float*>(Th));
        TNH_CHECK_LAUNCH("narrow");
        GemmArgs g1{static_cast<int>(dd * B), static_cast<int>(P * wl), static_cast<int>(Q), Th,
                    dd * B, nth, LT, P * wl, sLT, X1, P * wl, nx1, 0,
                    live, live ? 1 : 0, static_cast<int>(B), wl};
        TNH_TRY((gemm_any<TC, TC, TC, TC>(g1, batch, 0, st, tile1)));
      }
    }
    if (Th == nullptr) {
      GemmArgs g{static_cast<int>(dd * B), static_cast<int>(P * wl), static_cast<int>(Q), theta,
                 dd * B, sth, LT, P * wl, sLT, X1, P * wl, nx1, 0,
                 live, live ? 1 : 0, static_cast<int>(B), wl};
      TNH_TRY((gemm_any<TC, TIO, TC, TC>(g, batch, 0, st, tile1)));
    }
  }
  if ((stages & 2) && !fused) {  // S-mix
    MixArgs m{B, P, static_cast<int>(dd * wr), wr, static_cast<int>(wl * dd), static_cast<int>(dd),
              P * wl, wl, 1, B * P * wl, P * dd, dd, 1, B * P * dd, sS, nx1, nx2};
    TNH_TRY((mix<TC, TC, TC>(m, batch, S, X1, X2, st)));
  }
  if (stages & 4) {  // GEMM3: M = p d1 d2, N = b', K = wr b
    GemmArgs g{static_cast<int>(P * dd), static_cast<int>(B), static_cast<int>(wr * B), X2, P * dd,
               nx2, RT, B, sRT, out, B, sout, 0,
               live, live ? 2 : 0, static_cast<int>(B), static_cast<int>(dd)};
    TNH_TRY((gemm_any<TC, TC, TC, TIO>(g, batch, 0, st, tile3)));
  }
  return TNH_OK;
}

// ---- environments ------------------------------------------------------------------------------

template <typename T>
int left_env(int batch, int a, int d, int c, int wl, int wr, const void* L, long sL,
             const void* A, long sA, const void* W, long sW, void* out, long sout, void* ws,
             size_t ws_bytes, hipStream_t st) {
  const long n1 = (long)wl * a * d * c, n2 = (long)a * d * c * wr, nm = (long)wl * wr * d * d;
  Arena ar(ws, ws_bytes);
  T* T1 = ar.take<T>(n1 * batch);
  T* T2 = ar.take<T>(n2 * batch);
  T* Mx = ar.take<T>(nm * batch);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "update_left_env: workspace %zu < %zu bytes", ws_bytes, ar.used);
  hipLaunchKernelGGL(build_wmix_kernel<T>, dim3(grid_for(nm), 1, batch), dim3(256), 0, st, 0, wl,
                     wr, d, static_cast<const T*>(W), sW, Mx, nm);
  TNH_CHECK_LAUNCH("build_wmix");
  {  // T1[(w q)][(s e)] = sum_p L[p][(w q)] conj(A)[p][(s e)]
    GemmArgs g{wl * a, d * c, a, L, (long)wl * a, sL, A, (long)d * c, sA, T1, (long)d * c, n1, 0};
    TNH_TRY((gemm_any<T, T, T, T>(g, batch, 1, st, 0)));
  }
  {  // T2[q][y][e][r] = sum_{w s} Mx[(y r)][(w s)] T1[w][q][s][e]; groups (q, e)
    MixArgs m{a, c, d * wr, wr, wl * d, d, (long)d * c, 1, (long)a * d * c, c,
              (long)d * c * wr, wr, (long)c * wr, 1, nm, n1, n2};
    TNH_TRY((mix<T, T, T>(m, batch, Mx, T1, T2, st)));
  }
  {  // L'[(e r)][f] = sum_{(q y)} T2[(q y)][(e r)] A[(q y)][f]
    GemmArgs g{c * wr, c, a * d, T2, (long)c * wr, n2, A, c, sA, out, c, sout, 0};
    TNH_TRY((gemm_any<T, T, T, T>(g, batch, 0, st, 0)));
  }
  return TNH_OK;
}

template <typename T>
int right_env(int batch, int c, int d, int b, int wl, int wr, const void* R, long sR,
              const void* B, long sB, const void* W, long sW, void* out, long sout, void* ws,
              size_t ws_bytes, hipStream_t st) {
  const long nb = (long)b * c * d, n1 = (long)wr * b * c * d, n2 = (long)d * b * c * wl;
  const long nm = (long)wl * wr * d * d;
  Arena ar(ws, ws_bytes);
  T* Bx = ar.take<T>(nb * batch);
  T* Bp = ar.take<T>(nb * batch);
  T* T1 = ar.take<T>(n1 * batch);
  T* T2 = ar.take<T>(n2 * batch);
  T* Mx = ar.take<T>(nm * batch);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "update_right_env: workspace %zu < %zu bytes", ws_bytes,
              ar.used);
  // B(c, d, b) strides (d*b, b, 1) for (u, s, x)
  TNH_TRY(permute3<T>(batch, b, c, d, 1, (long)d * b, b, B, sB, Bx, nb, st));  // Bx[x][u][s]
  TNH_TRY(permute3<T>(batch, d, b, c, b, 1, (long)d * b, B, sB, Bp, nb, st));  // Bp[p][y][v]
  hipLaunchKernelGGL(build_wmix_kernel<T>, dim3(grid_for(nm), 1, batch), dim3(256), 0, st, 1, wl,
                     wr, d, static_cast<const T*>(W), sW, Mx, nm);
  TNH_CHECK_LAUNCH("build_wmix");
  {  // T1[(r y)][(u s)] = sum_x R[x][(r y)] conj(Bx)[x][(u s)]
    GemmArgs g{wr * b, c * d, b, R, (long)wr * b, sR, Bx, (long)c * d, nb, T1, (long)c * d, n1, 0};
    TNH_TRY((gemm_any<T, T, T, T>(g, batch, 1, st, 0)));
  }
  {  // T2[p][y][u][w] = sum_{r s} Mx[(p w)][(r s)] T1[r][y][u][s]; groups (y, u)
    MixArgs m{b, c, d * wl, wl, wr * d, d, (long)c * d, d, (long)b * c * d, 1,
              (long)c * wl, wl, (long)b * c * wl, 1, nm, n1, n2};
    TNH_TRY((mix<T, T, T>(m, batch, Mx, T1, T2, st)));
  }
  {  // R'[(u w)][v] = sum_{(p y)} T2[(p y)][(u w)] Bp[(p y)][v]
    GemmArgs g{c * wl, c, d * b, T2, (long)c * wl, n2, Bp, c, nb, out, c, sout, 0};
    TNH_TRY((gemm_any<T, T, T, T>(g, batch, 0, st, 0)));
  }
  return TNH_OK;
}

// ---- backward: the adjoint matvec and the environment VJPs -------------------------------------

// The operator of H_eff^H in the layouts the matvec streams: LT'[q][p][w] = conj(L[q][w][p]),
// RT'[r][b][b'] = conj(R[b][r][b']), S' from W1, W2 with s and s' swapped, conjugated.
template <typename T>
int fused_adjoint(int a, int wl, int wm, int wr, int d1, int d2, int b, const void* L,
                  const void* W1, const void* W2, const void* R, T* LT, T* S, T* RT,
                  hipStream_t st) {
  TNH_TRY((permute3<T, true>(1, a, a, wl, (long)wl * a, 1, a, L, 0, LT, 0, st)));
  TNH_TRY((permute3<T, true>(1, wr, b, b, b, (long)wr * b, 1, R, 0, RT, 0, st)));
  const int total = d1 * d2 * wr * wl * d1 * d2;
  hipLaunchKernelGGL(build_s_adjoint_kernel<T>, dim3(grid_for(total)), dim3(256), 0, st, wl, wm,
                     wr, d1, d2, static_cast<const T*>(W1), static_cast<const T*>(W2), S);
  TNH_CHECK_LAUNCH("build_s_adjoint");
  return TNH_OK;
}

template <typename T>
int conj_inplace(long n, T* x, hipStream_t st) {
  if constexpr (real_of<T>::cplx) {
    if (n == 0) return TNH_OK;
    hipLaunchKernelGGL(conj_kernel<T>, dim3(grid_for(n)), dim3(256), 0, st, n, x);
    TNH_CHECK_LAUNCH("conj");
  }
  return TNH_OK;
}

inline size_t op_bytes(size_t e, int a, int d1, int d2, int b, int wl, int wr) {
  const size_t dd = (size_t)d1 * d2;
  return align_up((size_t)a * a * wl * e, 256) + align_up(dd * wr * wl * dd * e, 256) +
         align_up((size_t)wr * b * b * e, 256);
}

template <typename T>
int matvec2_bwd(int a, int d1, int d2, int b, int wl, int wm, int wr, const void* L,
                const void* W1, const void* W2, const void* R, const void* ybar, void* theta_bar,
                void* ws, size_t ws_bytes, hipStream_t st) {
  const size_t dd = (size_t)d1 * d2;
  Arena ar(ws, ws_bytes);
  T* LT = ar.take<T>((size_t)a * a * wl);
  T* S = ar.take<T>(dd * wr * wl * dd);
  T* RT = ar.take<T>((size_t)wr * b * b);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "tnh_matvec2_bwd: workspace %zu < %zu bytes", ws_bytes,
              ar.used);
  const size_t used = align_up(ar.used, 256);
  TNH_TRY(fused_adjoint<T>(a, wl, wm, wr, d1, d2, b, L, W1, W2, R, LT, S, RT, st));
  return apply_two_site<T, T>(1, a, d1, d2, b, wl, wr, LT, 0, S, 0, RT, 0, ybar, 0, theta_bar, 0,
                              static_cast<char*>(ws) + used, ws_bytes - used, st, 0, 0, 7);
}

// VJP of L' = update_left_env(L, A, W) with respect to L: Lbar = conj(R'(conj(Lbar'), A, W)),
// R' the right-environment update (A in the bra slot conjugated); for real data
// Lbar = R'(Lbar', A, W). The VJP of the right update is the mirror, through the left update.
template <typename T>
int env_bwd(bool left, int a, int d, int c, int wl, int wr, const void* gout, const void* A,
            const void* W, void* gin, void* ws, size_t ws_bytes, hipStream_t st) {
  // left: gout (c, wr, c), A (a, d, c), gin (a, wl, a); right: gout (a, wl, a), B (a, d, c),
  // gin (c, wr, c)
  const long ng = left ? (long)c * wr * c : (long)a * wl * a;
  const long ni = left ? (long)a * wl * a : (long)c * wr * c;
  Arena ar(ws, ws_bytes);
  T* g = ar.take<T>(ng);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "tnh_env_bwd: workspace %zu < %zu bytes", ws_bytes, ar.used);
  const size_t used = align_up(ar.used, 256);
  const void* gsrc = gout;
  if constexpr (real_of<T>::cplx) {
    TNH_TRY((permute3<T, true>(1, 1, 1, ng, 0, 0, 1, gout, 0, g, 0, st)));
    gsrc = g;
  }
  void* rest = static_cast<char*>(ws) + used;
  const size_t rest_bytes = ws_bytes - used;
  if (left) {
    TNH_TRY(right_env<T>(1, a, d, c, wl, wr, gsrc, 0, A, 0, W, 0, gin, 0, rest, rest_bytes, st));
  } else {
    TNH_TRY(left_env<T>(1, a, d, c, wl, wr, gsrc, 0, A, 0, W, 0, gin, 0, rest, rest_bytes, st));
  }
  return conj_inplace<T>(ni, static_cast<T*>(gin), st);
}

}  // namespace
}  // namespace tnh

using namespace tnh;

// The BF16 operator: LT and RT in bf16 (the two GEMMs' B operands), S in float (the MPO mix runs
// in float); L and R arrive in bf16, W1 and W2 in float.
int fused_two_site_bf16(int batch, int a, int wl, int wm, int wr, int d1, int d2, int b,
                        const void* L, long sL, const void* W1, long sW1, const void* W2, long sW2,
                        const void* R, long sR, void* LT, long sLT, void* S, long sS, void* RT,
                        long sRT, hipStream_t st) {
  TNH_TRY(permute3<__bf16>(batch, a, a, wl, 1, (long)wl * a, a, L, sL, LT, sLT, st));
  TNH_TRY(permute3<__bf16>(batch, wr, b, b, b, 1, (long)wr * b, R, sR, RT, sRT, st));
  const int total = d1 * d2 * wr * wl * d1 * d2;
  hipLaunchKernelGGL(build_s_kernel<float>, dim3(grid_for(total), 1, batch), dim3(256), 0, st, wl,
                     wm, wr, d1, d2, static_cast<const float*>(W1), sW1,
                     static_cast<const float*>(W2), sW2, static_cast<float*>(S), sS);
  TNH_CHECK_LAUNCH("build_s");
  return TNH_OK;
}

// The BF16 matvec: GEMM1 on bf16 matrix cores (theta rounded to bf16 on the way into LDS, float
// accumulation, X1 in float), the MPO mix in float writing X2 in bf16, GEMM3 on bf16 matrix cores
// into the io type.
template <typename TIO>
int apply_two_site_bf16(int batch, int a, int d1, int d2, int b, int wl, int wr, const void* LT,
                        long sLT, const void* S, long sS, const void* RT, long sRT,
                        const void* theta, long sth, void* out, long sout, void* ws,
                        size_t ws_bytes, hipStream_t st, int tile1, int tile3, int stages) {
  const long P = a, Q = a, B = b, dd = (long)d1 * d2;
  const long nx1 = dd * B * P * wl, nx2 = (long)wr * B * P * dd;
  Arena ar(ws, ws_bytes);
  float* X1 = ar.take<float>(nx1 * batch);
  __bf16* X2 = ar.take<__bf16>(nx2 * batch);
  TNH_REQUIRE(ar.ok(), TNH_ERR_WS, "apply_two_site_fused (bf16): workspace %zu < %zu bytes",
              ws_bytes, ar.used);
  if (stages & 1) {
    GemmArgs g{static_cast<int>(dd * B), static_cast<int>(P * wl), static_cast<int>(Q), theta,
               dd * B, sth, LT, P * wl, sLT, X1, P * wl, nx1, 0};
    TNH_TRY((launch_gemm_bf16<TIO, __bf16, float>(g, batch, st, tile1)));
  }
  if (stages & 2) {
    MixArgs m{B, P, static_cast<int>(dd * wr), wr, static_cast<int>(wl * dd), static_cast<int>(dd),
              P * wl, wl, 1, B * P * wl, P * dd, dd, 1, B * P * dd, sS, nx1, nx2};
    TNH_TRY((mix<float, float, __bf16>(m, batch, S, X1, X2, st)));
  }
  if (stages & 4) {
    GemmArgs g{static_cast<int>(P * dd), static_cast<int>(B), static_cast<int>(wr * B), X2, P * dd,
               nx2, RT, B, sRT, out, B, sout, 0};
    TNH_TRY((launch_gemm_bf16<__bf16, __bf16, TIO>(g, batch, st, tile3)));
  }
  return TNH_OK;
}

extern "C" {

int tnh_fused_two_site_batched(int dtype, int batch, int a, int wl, int wm, int wr, int d1, int d2,
                               int b, const void* L, long sL, const void* W1, long sW1,
                               const void* W2, long sW2, const void* R, long sR, void* LT,
                               long sLT, void* S, long sS, void* RT, long sRT,
                               tnh_stream_t stream) {
  TNH_REQUIRE(a > 0 && b > 0 && wl > 0 && wm > 0 && wr > 0 && d1 > 0 && d2 > 0, TNH_ERR_ARG,
              "tnh_fused_two_site: sizes must be positive");
  TNH_REQUIRE(batch >= 1 && batch <= 65535, TNH_ERR_ARG, "tnh_fused_two_site: batch %d", batch);
  hipStream_t st = static_cast<hipStream_t>(stream);
#define TNH_FTS(T) \
  fused_two_site<T>(batch, a, wl, wm, wr, d1, d2, b, L, sL, W1, sW1, W2, sW2, R, sR, LT, sLT, S, sS, RT, sRT, st)
  switch (dtype) {
    case TNH_F64: return TNH_FTS(double);
    case TNH_F32: return TNH_FTS(float);
    case TNH_C128: return TNH_FTS(c128);
    case TNH_C64: return TNH_FTS(c64);
    case TNH_BF16:
      return fused_two_site_bf16(batch, a, wl, wm, wr, d1, d2, b, L, sL, W1, sW1, W2, sW2, R, sR,
                                 LT, sLT, S, sS, RT, sRT, st);
    default: set_error("tnh_fused_two_site: unsupported dtype %d", dtype); return TNH_ERR_UNSUPPORTED;
  }
#undef TNH_FTS
}

int tnh_fused_two_site(int dtype, int a, int wl, int wm, int wr, int d1, int d2, int b,
                       const void* L, const void* W1, const void* W2, const void* R, void* LT,
                       void* S, void* RT, tnh_stream_t stream) {
  return tnh_fused_two_site_batched(dtype, 1, a, wl, wm, wr, d1, d2, b, L, 0, W1, 0, W2, 0, R, 0,
                                    LT, 0, S, 0, RT, 0, stream);
}

size_t tnh_two_site_ws_bytes(int dtype, int a, int d1, int d2, int b, int wl, int wr) {
  const size_t dd = (size_t)d1 * d2;
  if (dtype == TNH_BF16)  // X1 in float, X2 in bf16
    return align_up(dd * b * a * wl * 4, 256) + align_up((size_t)wr * b * a * dd * 2, 256) + 256;
  const size_t e = dsize(dtype);
  // float32 / complex64: room for theta narrowed from float64 / complex128 vectors
  const size_t th = (dtype == TNH_F32 || dtype == TNH_C64) ? align_up((size_t)a * dd * b * e, 256)
                                                           : 0;
  return align_up(dd * b * a * wl * e, 256) + align_up((size_t)wr * b * a * dd * e, 256) + th +
         256;
}

int tnh_apply_two_site_batched(int dtype, int io_dtype, int batch, int a, int d1, int d2, int b,
                               int wl, int wr, const void* LT, long sLT, const void* S, long sS,
                               const void* RT, long sRT, const void* theta, long sth, void* out,
                               long sout, void* ws, size_t ws_bytes, int tile1, int tile3,
                               int stages, tnh_stream_t stream) {
  TNH_REQUIRE(a > 0 && b > 0 && wl > 0 && wr > 0 && d1 > 0 && d2 > 0, TNH_ERR_ARG,
              "tnh_apply_two_site_fused: sizes must be positive");
  TNH_REQUIRE(stages >= 1 && stages <= 7, TNH_ERR_ARG, "tnh_apply_two_site: stages %d", stages);
  TNH_REQUIRE(batch >= 1 && batch <= 65535, TNH_ERR_ARG, "tnh_apply_two_site: batch %d", batch);
  hipStream_t st = static_cast<hipStream_t>(stream);
#define TNH_A2S(TC, TIO)                                                                       \
  apply_two_site<TC, TIO>(batch, a, d1, d2, b, wl, wr, LT, sLT, S, sS, RT, sRT, theta, sth, out, \
                          sout, ws, ws_bytes, st, tile1, tile3, stages)
  if (dtype == TNH_F64 && io_dtype == TNH_F64) return TNH_A2S(double, double);
  if (dtype == TNH_F32 && io_dtype == TNH_F32) return TNH_A2S(float, float);
  if (dtype == TNH_F32 && io_dtype == TNH_F64) return TNH_A2S(float, double);
  if (dtype == TNH_C128 && io_dtype == TNH_C128) return TNH_A2S(c128, c128);
  if (dtype == TNH_C64 && io_dtype == TNH_C64) return TNH_A2S(c64, c64);
  if (dtype == TNH_C64 && io_dtype == TNH_C128) return TNH_A2S(c64, c128);
#undef TNH_A2S
  if (dtype == TNH_BF16 && io_dtype == TNH_F64)
    return apply_two_site_bf16<double>(batch, a, d1, d2, b, wl, wr, LT, sLT, S, sS, RT, sRT,
                                       theta, sth, out, sout, ws, ws_bytes, st, tile1, tile3,
                                       stages);
  set_error("tnh_apply_two_site_fused: unsupported dtype pair (%d, %d)", dtype, io_dtype);
  return TNH_ERR_UNSUPPORTED;
}

int tnh_apply_two_site_ragged(int dtype, int io_dtype, int batch, int a, int d1, int d2, int b,
                              int wl, int wr, const void* LT, long sLT, const void* S, long sS,
                              const void* RT, long sRT, const void* theta, long sth, void* out,
                              long sout, const int* live, void* ws, size_t ws_bytes,
                              tnh_stream_t stream) {
  TNH_REQUIRE(a > 0 && b > 0 && wl > 0 && wr > 0 && d1 > 0 && d2 > 0 && live, TNH_ERR_ARG,
              "tnh_apply_two_site_ragged: sizes must be positive and live given");
  TNH_REQUIRE(batch >= 1 && batch <= 65535, TNH_ERR_ARG, "tnh_apply_two_site_ragged: batch %d",
              batch);
  hipStream_t st = static_cast<hipStream_t>(stream);
  if (dtype == TNH_C128 && io_dtype == TNH_C128)
    return apply_two_site<c128, c128>(batch, a, d1, d2, b, wl, wr, LT, sLT, S, sS, RT, sRT, theta,
                                      sth, out, sout, ws, ws_bytes, st, 0, 0, 7, live);
  if (dtype == TNH_C64 && io_dtype == TNH_C128)
    return apply_two_site<c64, c128>(batch, a, d1, d2, b, wl, wr, LT, sLT, S, sS, RT, sRT, theta,
                                     sth, out, sout, ws, ws_bytes, st, 0, 0, 7, live);
  set_error("tnh_apply_two_site_ragged: complex128 vectors with a complex128 or complex64 "
            "operator only (dtype pair %d, %d)", dtype, io_dtype);
  return TNH_ERR_UNSUPPORTED;
}

int tnh_apply_two_site_stages(int dtype, int io_dtype, int a, int d1, int d2, int b, int wl,
                              int wr, const void* LT, const void* S, const void* RT,
                              const void* theta, void* out, void* ws, size_t ws_bytes, int tile1,
                              int tile3, int stages, tnh_stream_t stream) {
  return tnh_apply_two_site_batched(dtype, io_dtype, 1, a, d1, d2, b, wl, wr, LT, 0, S, 0, RT, 0,
                                    theta, 0, out, 0, ws, ws_bytes, tile1, tile3, stages, stream);
}

int tnh_apply_two_site_fused(int dtype, int io_dtype, int a, int d1, int d2, int b, int wl, int wr,
                             const void* LT, const void* S, const void* RT, const void* theta,
                             void* out, void* ws, size_t ws_bytes, int tile1, int tile3,
                             tnh_stream_t stream) {
  return tnh_apply_two_site_batched(dtype, io_dtype, 1, a, d1, d2, b, wl, wr, LT, 0, S, 0, RT, 0,
                                    theta, 0, out, 0, ws, ws_bytes, tile1, tile3, 7, stream);
}

size_t tnh_env_ws_bytes(int dtype, int a, int d, int c, int wl, int wr) {
  // the larger of the left (bonds a -> c) and right (bonds a <- c) workspaces
  const size_t e = dsize(dtype);
  const size_t w = (size_t)wl > (size_t)wr ? wl : wr;
  const size_t big = (size_t)a * d * c;
  return 2 * align_up(big * e, 256) + 2 * align_up(big * w * e, 256) +
         align_up((size_t)wl * wr * d * d * e, 256) + 256;
}

int tnh_update_left_env_batched(int dtype, int batch, int a, int d, int c, int wl, int wr,
                                const void* L, long sL, const void* A, long sA, const void* W,
                                long sW, void* out, long sout, void* ws, size_t ws_bytes,
                                tnh_stream_t stream) {
  TNH_REQUIRE(a > 0 && c > 0 && d > 0 && wl > 0 && wr > 0, TNH_ERR_ARG,
              "tnh_update_left_env: sizes must be positive");
  TNH_REQUIRE(batch >= 1 && batch <= 65535, TNH_ERR_ARG, "tnh_update_left_env: batch %d", batch);
  hipStream_t st = static_cast<hipStream_t>(stream);
#define TNH_LE(T) left_env<T>(batch, a, d, c, wl, wr, L, sL, A, sA, W, sW, out, sout, ws, ws_bytes, st)
  switch (dtype) {
    case TNH_F64: return TNH_LE(double);
    case TNH_F32: return TNH_LE(float);
    case TNH_C128: return TNH_LE(c128);
    case TNH_C64: return TNH_LE(c64);
    default: set_error("tnh_update_left_env: unsupported dtype %d", dtype); return TNH_ERR_UNSUPPORTED;
  }
#undef TNH_LE
}

int tnh_update_left_env(int dtype, int a, int d, int c, int wl, int wr, const void* L,
                        const void* A, const void* W, void* out, void* ws, size_t ws_bytes,
                        tnh_stream_t stream) {
  return tnh_update_left_env_batched(dtype, 1, a, d, c, wl, wr, L, 0, A, 0, W, 0, out, 0, ws,
                                     ws_bytes, stream);
}

int tnh_update_right_env_batched(int dtype, int batch, int c, int d, int b, int wl, int wr,
                                 const void* R, long sR, const void* B, long sB, const void* W,
                                 long sW, void* out, long sout, void* ws, size_t ws_bytes,
                                 tnh_stream_t stream) {
  TNH_REQUIRE(b > 0 && c > 0 && d > 0 && wl > 0 && wr > 0, TNH_ERR_ARG,
              "tnh_update_right_env: sizes must be positive");
  TNH_REQUIRE(batch >= 1 && batch <= 65535, TNH_ERR_ARG, "tnh_update_right_env: batch %d", batch);
  hipStream_t st = static_cast<hipStream_t>(stream);
#define TNH_RE(T) right_env<T>(batch, c, d, b, wl, wr, R, sR, B, sB, W, sW, out, sout, ws, ws_bytes, st)
  switch (dtype) {
    case TNH_F64: return TNH_RE(double);
    case TNH_F32: return TNH_RE(float);
    case TNH_C128: return TNH_RE(c128);
    case TNH_C64: return TNH_RE(c64);
    default: set_error("tnh_update_right_env: unsupported dtype %d", dtype); return TNH_ERR_UNSUPPORTED;
  }
#undef TNH_RE
}

int tnh_update_right_env(int dtype, int c, int d, int b, int wl, int wr, const void* R,
                         const void* B, const void* W, void* out, void* ws, size_t ws_bytes,
                         tnh_stream_t stream) {
  return tnh_update_right_env_batched(dtype, 1, c, d, b, wl, wr, R, 0, B, 0, W, 0, out, 0, ws,
                                      ws_bytes, stream);
}

size_t tnh_matvec2_bwd_ws_bytes(int dtype, int a, int d1, int d2, int b, int wl, int wr) {
  return op_bytes(dsize(dtype), a, d1, d2, b, wl, wr) + 256 +
         tnh_two_site_ws_bytes(dtype, a, d1, d2, b, wl, wr);
}

int tnh_matvec2_bwd(int dtype, int a, int d1, int d2, int b, int wl, int wm, int wr,
                    const void* L, const void* W1, const void* W2, const void* R,
                    const void* ybar, void* theta_bar, void* ws, size_t ws_bytes,
                    tnh_stream_t stream) {
  TNH_REQUIRE(a > 0 && b > 0 && wl > 0 && wm > 0 && wr > 0 && d1 > 0 && d2 > 0, TNH_ERR_ARG,
              "tnh_matvec2_bwd: sizes must be positive");
  hipStream_t st = static_cast<hipStream_t>(stream);
#define TNH_BWD(T) \
  matvec2_bwd<T>(a, d1, d2, b, wl, wm, wr, L, W1, W2, R, ybar, theta_bar, ws, ws_bytes, st)
  switch (dtype) {
    case TNH_F64: return TNH_BWD(double);
    case TNH_F32: return TNH_BWD(float);
    case TNH_C128: return TNH_BWD(c128);
    case TNH_C64: return TNH_BWD(c64);
  }
#undef TNH_BWD
  set_error("tnh_matvec2_bwd: unsupported dtype %d", dtype);
  return TNH_ERR_UNSUPPORTED;
}

size_t tnh_env_bwd_ws_bytes(int dtype, int a, int d, int c, int wl, int wr) {
  const size_t e = dsize(dtype);
  const size_t big = (size_t)a * a * wl > (size_t)c * c * wr ? (size_t)a * a * wl
                                                             : (size_t)c * c * wr;
  return align_up(big * e, 256) + 256 + tnh_env_ws_bytes(dtype, a, d, c, wl, wr);
}

static int env_bwd_any(bool left, int dtype, int a, int d, int c, int wl, int wr,
                       const void* gout, const void* A, const void* W, void* gin, void* ws,
                       size_t ws_bytes, tnh_stream_t stream) {
  TNH_REQUIRE(a > 0 && c > 0 && d > 0 && wl > 0 && wr > 0, TNH_ERR_ARG,
              "tnh_env_bwd: sizes must be positive");
  hipStream_t st = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case TNH_F64: return env_bwd<double>(left, a, d, c, wl, wr, gout, A, W, gin, ws, ws_bytes, st);
    case TNH_F32: return env_bwd<float>(left, a, d, c, wl, wr, gout, A, W, gin, ws, ws_bytes, st);
    case TNH_C128: return env_bwd<c128>(left, a, d, c, wl, wr, gout, A, W, gin, ws, ws_bytes, st);
    case TNH_C64: return env_bwd<c64>(left, a, d, c, wl, wr, gout, A, W, gin, ws, ws_bytes, st);
  }
  set_error("tnh_env_bwd: unsupported dtype %d", dtype);
  return TNH_ERR_UNSUPPORTED;
}

int tnh_env_left_bwd(int dtype, int a, int d, int c, int wl, int wr, const void* Lbar_out,
                     const void* A, const void* W, void* Lbar, void* ws, size_t ws_bytes,
                     tnh_stream_t stream) {
  return env_bwd_any(true, dtype, a, d, c, wl, wr, Lbar_out, A, W, Lbar, ws, ws_bytes, stream);
}

int tnh_env_right_bwd(int dtype, int c, int d, int b, int wl, int wr, const void* Rbar_out,
                      const void* B, const void* W, void* Rbar, void* ws, size_t ws_bytes,
                      tnh_stream_t stream) {
  return env_bwd_any(false, dtype, c, d, b, wl, wr, Rbar_out, B, W, Rbar, ws, ws_bytes, stream);
}

}  // extern "C"
