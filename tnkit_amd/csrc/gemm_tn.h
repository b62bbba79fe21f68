// K-major ("TN") GEMM on CDNA4 matrix cores:  C[m][n] = sum_k A[k][m] * B[k][n]
//
// Every large contraction of the two-site path is written so that both operands are stored
// with the contracted index outermost (K rows of contiguous M or N entries). Tiles then move
// from HBM to LDS as plain row copies, and the MFMA fragments are read from LDS as 16
// consecutive entries per k row, which is bank-conflict free with the 16-element row pad.
//
// MFMA shapes (wave64):
//   double : v_mfma_f64_16x16x4f64   lane l holds A[l%16][l/16], B[l/16][l%16]
//   float  : v_mfma_f32_16x16x4f32   same operand layout
//   result : float  -- lane l, item r holds D[4*(l/16) + r][l%16]
//            double -- lane l, item r holds D[4*r + l/16][l%16]   (rows interleaved)
//
// Storage types may differ from the compute type (A or B stored in double and rounded to
// float on the way into LDS, C stored in double from a float accumulator), which is how the
// mixed-precision matvec takes float64 Krylov vectors without a separate conversion pass.
#pragma once
#include <type_traits>

#include "cplx.h"
#include "tnh_common.h"

namespace tnh {

template <typename T>
struct Mfma;

template <>
struct Mfma<double> {
  using acc_t = d4;
  static __device__ __forceinline__ acc_t run(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  // row of result item r held by a lane in row-group kq = lane / 16
  static __device__ __forceinline__ int row(int kq, int r) { return 4 * r + kq; }
};

template <>
struct Mfma<float> {
  using acc_t = f4;
  static __device__ __forceinline__ acc_t run(float a, float b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int kq, int r) { return 4 * kq + r; }
};

struct GemmArgs {
  int M, N, K;
  const void* A;
  long lda;
  long strideA;  // batch strides (elements)
  const void* B;
  long ldb;
  long strideB;
  void* C;
  long ldc;
  long strideC;
  int accumulate;  // C += A^T B instead of C = A^T B
  // Live extents of batch entries (the complex direct-to-LDS kernel only; others ignore them):
  // live[2z], live[2z+1] = (x, y) of entry z, and the product outside its live block is zero by
  // construction of the operands. live_mode 1 (the two-site GEMM1): a tile of rows m0 with
  // m0 mod lp1 >= y or of columns n0 >= x lp2 is dead, K stops at x. live_mode 2 (GEMM3): rows
  // m0 >= x lp2 or columns n0 >= y are dead, and of K only rows k with k mod lp1 < y are read.
  // Dead tiles store zeros (or leave C as it is when accumulating).
  const int* live = nullptr;
  int live_mode = 0;
  int lp1 = 0, lp2 = 0;
};

// Tile -> block mapping. Workgroups are dispatched round-robin over the 8 XCDs, each with its
// own L2; consecutive logical tiles are remapped onto one XCD, and tiles are walked in groups
// of `kGroup` tile rows so that neighbours share A and B panels in that XCD's L2.
__device__ __forceinline__ void tile_of(int bid, int nbm, int nbn, int& tm, int& tn) {
  constexpr int kXcd = 8;
  constexpr int kGroup = 8;
  const int total = nbm * nbn;
  int pid = bid;
  if (total % kXcd == 0) pid = (bid % kXcd) * (total / kXcd) + bid / kXcd;
  const int per_group = kGroup * nbn;
  const int g = pid / per_group;
  const int first = g * kGroup;
  const int gsize = min(nbm - first, kGroup);
  const int in = pid - g * per_group;
  tm = first + in % gsize;
  tn = in / gsize;
}

// One thread's share of a K-major tile: BK rows of BMN contiguous entries, moved as vectors of V
// elements (V = 1 or 16 bytes' worth when the leading dimension and base allow it).
template <typename TC, typename TS, int BK, int BMN, int NT, int V>
struct TileLoader {
  static constexpr int NV = BK * BMN / V / NT;  // vectors per thread
  static_assert((BK * BMN) % (V * NT) == 0, "tile not divisible by threads");
  TC r[NV * V];
  __device__ __forceinline__ void load(const TS* __restrict__ src, long ld, int k0, int K, int c0,
                                       int C, int t) {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int e = t + i * NT;
      const int kr = e / (BMN / V), cc = (e % (BMN / V)) * V;
      const int k = k0 + kr, c = c0 + cc;
      if constexpr (V == 1) {
        r[i] = (k < K && c < C) ? static_cast<TC>(src[(long)k * ld + c]) : TC(0);
      } else {
        typedef TS vec_t __attribute__((ext_vector_type(V)));
        if (k < K && c < C) {  // C % V == 0 and c % V == 0, so the whole vector is in bounds
          const vec_t x = *reinterpret_cast<const vec_t*>(src + (long)k * ld + c);
#pragma unroll
          for (int v = 0; v < V; ++v) r[i * V + v] = static_cast<TC>(x[v]);
        } else {
#pragma unroll
          for (int v = 0; v < V; ++v) r[i * V + v] = TC(0);
        }
      }
    }
  }
  __device__ __forceinline__ void store(TC* __restrict__ dst, int lds, int t) const {
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int e = t + i * NT;
      const int kr = e / (BMN / V), cc = (e % (BMN / V)) * V;
#pragma unroll
      for (int v = 0; v < V; ++v) dst[kr * lds + cc + v] = r[i * V + v];
    }
  }
};

template <typename TC, typename TA, typename TB, typename TO, int BM, int BN, int BK, int WM, int WN,
          int VA, int VB>
__global__ void __launch_bounds__(WM* WN * 64)
    gemm_tn_kernel(GemmArgs g) {
  using MF = Mfma<TC>;
  using acc_t = typename MF::acc_t;
  constexpr int NT = WM * WN * 64;
  constexpr int PAD = 16;
  constexpr int LDA_S = BM + PAD;
  constexpr int LDB_S = BN + PAD;
  constexpr int FM = BM / WM / 16;
  constexpr int FN = BN / WN / 16;
  static_assert(BK % 4 == 0, "BK must be a multiple of the MFMA k=4");
  static_assert(FM >= 1 && FN >= 1, "wave tile too small");

  __shared__ TC As[2][BK * LDA_S];
  __shared__ TC Bs[2][BK * LDB_S];

  const int M = g.M, N = g.N, K = g.K;
  const int nbm = (M + BM - 1) / BM;
  const int nbn = (N + BN - 1) / BN;
  int tm, tn;
  tile_of(blockIdx.x, nbm, nbn, tm, tn);
  const int m0 = tm * BM;
  const int n0 = tn * BN;
  const long bz = blockIdx.y;
  const TA* __restrict__ A = static_cast<const TA*>(g.A) + bz * g.strideA;
  const TB* __restrict__ B = static_cast<const TB*>(g.B) + bz * g.strideB;
  TO* __restrict__ C = static_cast<TO*>(g.C) + bz * g.strideC;

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave / WN;
  const int wn = wave % WN;
  const int row16 = lane & 15;
  const int kq = lane >> 4;

  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};

  TileLoader<TC, TA, BK, BM, NT, VA> la;
  TileLoader<TC, TB, BK, BN, NT, VB> lb;

  const int nk = (K + BK - 1) / BK;
  if (nk > 0) {
    la.load(A, g.lda, 0, K, m0, M, t);
    lb.load(B, g.ldb, 0, K, n0, N, t);
    la.store(As[0], LDA_S, t);
    lb.store(Bs[0], LDB_S, t);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) {
      la.load(A, g.lda, (kt + 1) * BK, K, m0, M, t);
      lb.load(B, g.ldb, (kt + 1) * BK, K, n0, N, t);
    }
    const TC* as = As[cur];
    const TC* bs = Bs[cur];
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      TC af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = as[(kk + kq) * LDA_S + wm * (BM / WM) + i * 16 + row16];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = bs[(kk + kq) * LDB_S + wn * (BN / WN) + j * 16 + row16];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MF::run(af[i], bf[j], acc[i][j]);
    }
    if (kt + 1 < nk) {
      la.store(As[cur ^ 1], LDA_S, t);
      lb.store(Bs[cur ^ 1], LDB_S, t);
    }
    __syncthreads();
  }

#pragma unroll
  for (int i = 0; i < FM; ++i) {
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int n = n0 + wn * (BN / WN) + j * 16 + row16;
      if (n >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * (BM / WM) + i * 16 + MF::row(kq, r);
        if (m < M) {
          TO* dst = C + (long)m * g.ldc + n;
          if (g.accumulate)
            *dst = static_cast<TO>(static_cast<TC>(*dst) + acc[i][j][r]);
          else
            *dst = static_cast<TO>(acc[i][j][r]);
        }
      }
    }
  }
}

// ---- direct-to-LDS variant --------------------------------------------------------------------
// Float or double operands stored in the compute type (no conversion on the way in). Tiles go
// from HBM straight into LDS with global_load_lds_dwordx4 (lane i of a wave writes bytes
// [16 i, 16 i + 16) after the wave's LDS base), so no VGPRs stage them and no ds_write is issued.
// Three LDS stages of BK rows form a ring: the tile two steps ahead is in flight while one is
// computed. Stages are separate arrays so that the compiler can tell a DMA into one stage from
// fragment reads of another and inserts no wait between them; the waits on the tile being
// consumed are explicit vmcnt counts before the workgroup barrier. Rows are unpadded (the DMA
// writes contiguous 1 KB per wave) and rotated by 16 elements per (row mod 4), so the four k rows
// of one fragment read fall in distinct bank ranges. 128 x 128 tiles, 2 x WNC waves of
// 64 x (128 / WNC) (WNC = 2: 4 waves of 64 x 64; WNC = 4: 8 waves of 64 x 32, two per SIMD).
// Requires M, N % 128 == 0, K % BK == 0, 16-byte aligned operands with lda, ldb and batch strides
// multiples of a 16-byte chunk (the launcher checks).
typedef __attribute__((address_space(3))) void* lds_void_t;

template <typename T, int BK>
struct DmaStages {
  T a0[BK * 128], a1[BK * 128], a2[BK * 128];
  T b0[BK * 128], b1[BK * 128], b2[BK * 128];
};

template <int S, typename SM>
__device__ __forceinline__ auto* dma_stage_a(SM& m) {
  if constexpr (S == 0) return m.a0;
  else if constexpr (S == 1) return m.a1;
  else return m.a2;
}

template <int S, typename SM>
__device__ __forceinline__ auto* dma_stage_b(SM& m) {
  if constexpr (S == 0) return m.b0;
  else if constexpr (S == 1) return m.b1;
  else return m.b2;
}

// s_waitcnt vmcnt(NW), the other counters left alone (gfx9 encoding)
template <int NW>
__device__ __forceinline__ void wait_vmcnt() {
  static_assert(NW >= 0 && NW < 64, "vmcnt range");
  __builtin_amdgcn_s_waitcnt((NW & 15) | ((NW >> 4) << 14) | (0x7 << 4) | (0xF << 8));
}

template <typename T, typename TO, int BK, int WNC = 2>
__global__ void __launch_bounds__(WNC * 128) gemm_dma_kernel(GemmArgs g) {
  using MF = Mfma<T>;
  using acc_t = typename MF::acc_t;
  constexpr int NT = WNC * 128;        // 2 x WNC waves of 64 x (128 / WNC)
  constexpr int BM = 128, FM = 4, FN = 128 / WNC / 16, WCOL = 128 / WNC;
  constexpr int EPC = 16 / sizeof(T);  // elements per 16-byte chunk
  constexpr int CPR = BM / EPC;        // chunks per tile row
  constexpr int NJ = BK * CPR / NT;    // chunks per thread per operand tile
  constexpr int ROT = 16 / EPC;        // the 16-element row rotation, in chunks
  static_assert(NJ * NT == BK * CPR && BK % 4 == 0, "dma tile shape");
  __shared__ __align__(16) DmaStages<T, BK> sm;
  const int M = g.M, N = g.N, K = g.K;
  int tm, tn;
  tile_of(blockIdx.x, M / BM, N / BM, tm, tn);
  const int m0 = tm * BM, n0 = tn * BM;
  const long bz = blockIdx.y;
  const T* __restrict__ A = static_cast<const T*>(g.A) + bz * g.strideA;
  const T* __restrict__ B = static_cast<const T*>(g.B) + bz * g.strideB;
  TO* __restrict__ C = static_cast<TO*>(g.C) + bz * g.strideC;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave / WNC, wn = wave % WNC, row16 = lane & 15, kq = lane >> 4;
  // element offsets, within a tile, of this thread's chunks (chunk c of the tile lands at LDS
  // position c; it holds the chunk of its row rotated back by ROT * (row mod 4))
  long goffA[NJ], goffB[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int c = j * NT + t;
    const int r = c / CPR, p = c % CPR;
    const int q = (p - ROT * (r & 3)) & (CPR - 1);
    goffA[j] = (long)r * g.lda + m0 + EPC * q;
    goffB[j] = (long)r * g.ldb + n0 + EPC * q;
  }
  acc_t acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};
  const int nk = K / BK;
  const int aoff = wm * 64 + row16 + 16 * kq, boff = wn * WCOL + row16 + 16 * kq;
  auto issue = [&](int kt, T* as, T* bs) {
    const T* a = A + (long)kt * BK * g.lda;
    const T* b = B + (long)kt * BK * g.ldb;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      __builtin_amdgcn_global_load_lds(static_cast<const void*>(a + goffA[j]),
                                       (lds_void_t)(as + (j * NT + wave * 64) * EPC), 16, 0, 0);
      __builtin_amdgcn_global_load_lds(static_cast<const void*>(b + goffB[j]),
                                       (lds_void_t)(bs + (j * NT + wave * 64) * EPC), 16, 0, 0);
    }
  };
  auto compute = [&](const T* as, const T* bs) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = as[(kk + kq) * BM + ((aoff + i * 16) & (BM - 1))];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = bs[(kk + kq) * BM + ((boff + j * 16) & (BM - 1))];
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) acc[i][j] = MF::run(af[i], bf[j], acc[i][j]);
    }
  };
  // one step of the ring on stage S: tile kt landed (at most the next tile's loads still in
  // flight), barrier, prefetch tile kt + 2 into the stage read two steps ago, compute tile kt
  auto step = [&](auto stage, int kt) {
    constexpr int S = decltype(stage)::value;
    if (kt + 1 < nk)
      wait_vmcnt<2 * NJ>();
    else
      wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    if (kt + 2 < nk) issue(kt + 2, dma_stage_a<(S + 2) % 3>(sm), dma_stage_b<(S + 2) % 3>(sm));
    compute(dma_stage_a<S>(sm), dma_stage_b<S>(sm));
  };
  if (nk > 0) issue(0, sm.a0, sm.b0);
  if (nk > 1) issue(1, sm.a1, sm.b1);
  for (int kt = 0; kt < nk; kt += 3) {
    step(std::integral_constant<int, 0>{}, kt);
    if (kt + 1 < nk) step(std::integral_constant<int, 1>{}, kt + 1);
    if (kt + 2 < nk) step(std::integral_constant<int, 2>{}, kt + 2);
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int n = n0 + wn * WCOL + j * 16 + row16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 64 + i * 16 + MF::row(kq, r);
        TO* dst = C + (long)m * g.ldc + n;
        if (g.accumulate)
          *dst = static_cast<TO>(static_cast<T>(*dst) + acc[i][j][r]);
        else
          *dst = static_cast<TO>(acc[i][j][r]);
      }
    }
}

// Complex form of gemm_dma_kernel (complex128 on float64 MFMAs, complex64 on float32 ones): the
// same three-stage direct-to-LDS ring and 16-element row rotation over the interleaved (re, im)
// elements as stored, and four real MFMAs per fragment product (re += ar br - ai bi,
// im += ar bi + ai br). A complex tile of BK rows carries twice the flops per LDS byte of the
// real one. 128 x 128 tiles, 8 waves of 64 x 32 (two per SIMD); C in TOR's complex type (the
// complex64 operator's last GEMM writes complex128). Requires M, N % 128 == 0, K % BK == 0 and
// 16-byte chunks (the launcher checks).
template <typename TR>
struct alignas(2 * sizeof(TR)) cpod {
  TR re, im;
};

template <typename TR, typename TOR, int BK>
__global__ void __launch_bounds__(512) gemm_dma_cplx_kernel(GemmArgs g, int conj_b) {
  using MF = Mfma<TR>;
  using acc_t = typename MF::acc_t;
  using T = cpod<TR>;
  using TO = cpod<TOR>;
  constexpr int WNC = 4, NT = 512;
  constexpr int BM = 128, FM = 4, FN = 2, WCOL = 32;
  constexpr int EPC = 16 / sizeof(T);  // complex elements per 16-byte chunk
  constexpr int CPR = BM / EPC;        // chunks per tile row
  constexpr int NJ = BK * CPR / NT;    // chunks per thread per operand tile
  constexpr int ROT = 16 / EPC;        // the 16-element row rotation, in chunks
  static_assert(NJ * NT == BK * CPR && BK % 4 == 0, "dma tile shape");
  __shared__ __align__(16) DmaStages<T, BK> sm;
  const int M = g.M, N = g.N, K = g.K;
  int tm, tn;
  tile_of(blockIdx.x, M / BM, N / BM, tm, tn);
  const int m0 = tm * BM, n0 = tn * BM;
  const long bz = blockIdx.y;
  const T* __restrict__ A = static_cast<const T*>(g.A) + bz * g.strideA;
  const T* __restrict__ B = static_cast<const T*>(g.B) + bz * g.strideB;
  TO* __restrict__ C = static_cast<TO*>(g.C) + bz * g.strideC;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave / WNC, wn = wave % WNC, row16 = lane & 15, kq = lane >> 4;
  const TR bsign = conj_b ? TR(-1) : TR(1);
  long goffA[NJ], goffB[NJ];
#pragma unroll
  for (int j = 0; j < NJ; ++j) {
    const int c = j * NT + t;
    const int r = c / CPR, p = c % CPR;
    const int q = (p - ROT * (r & 3)) & (CPR - 1);
    goffA[j] = (long)r * g.lda + m0 + EPC * q;
    goffB[j] = (long)r * g.ldb + n0 + EPC * q;
  }
  acc_t cr[FM][FN], ci[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) cr[i][j] = ci[i][j] = acc_t{};
  // K steps: all of them, or the live ones of a ragged batch entry (runs of spr steps every kmod
  // rows; a step is read whole, its rows past the live extent are zeros)
  int nk = K / BK, spr = nk, kstride = nk;
  if (g.live != nullptr && g.live_mode != 0) {
    const int x = g.live[2 * bz], y = g.live[2 * bz + 1];
    bool dead;
    int klim, kmod;
    if (g.live_mode == 1) {
      dead = (m0 % g.lp1) >= y || n0 >= x * g.lp2;
      klim = x;
      kmod = K;
    } else {
      dead = m0 >= x * g.lp2 || n0 >= y;
      klim = y;
      kmod = g.lp1;
    }
    if (dead) {
      if (!g.accumulate) {
        const TO z{TOR(0), TOR(0)};
        for (int e = t; e < BM * BM; e += NT) C[(long)(m0 + e / BM) * g.ldc + n0 + e % BM] = z;
      }
      return;
    }
    kstride = kmod / BK;
    spr = (min(klim, kmod) + BK - 1) / BK;
    nk = (K / kmod) * spr;
  }
  auto kphys = [&](int i) { return (i / spr) * kstride + i % spr; };
  const int aoff = wm * 64 + row16 + 16 * kq, boff = wn * WCOL + row16 + 16 * kq;
  auto issue = [&](int i, T* as, T* bs) {
    const int kt = kphys(i);
    const T* a = A + (long)kt * BK * g.lda;
    const T* b = B + (long)kt * BK * g.ldb;
#pragma unroll
    for (int j = 0; j < NJ; ++j) {
      __builtin_amdgcn_global_load_lds(static_cast<const void*>(a + goffA[j]),
                                       (lds_void_t)(as + (j * NT + wave * 64) * EPC), 16, 0, 0);
      __builtin_amdgcn_global_load_lds(static_cast<const void*>(b + goffB[j]),
                                       (lds_void_t)(bs + (j * NT + wave * 64) * EPC), 16, 0, 0);
    }
  };
  auto compute = [&](const T* as, const T* bs) {
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      T af[FM], bf[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) af[i] = as[(kk + kq) * BM + ((aoff + i * 16) & (BM - 1))];
#pragma unroll
      for (int j = 0; j < FN; ++j) bf[j] = bs[(kk + kq) * BM + ((boff + j * 16) & (BM - 1))];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const TR ar = af[i].re, ai = af[i].im, nai = -ai;
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          const TR br = bf[j].re, bi = bsign * bf[j].im;
          cr[i][j] = MF::run(ar, br, cr[i][j]);
          ci[i][j] = MF::run(ar, bi, ci[i][j]);
          cr[i][j] = MF::run(nai, bi, cr[i][j]);
          ci[i][j] = MF::run(ai, br, ci[i][j]);
        }
      }
    }
  };
  auto step = [&](auto stage, int kt) {
    constexpr int S = decltype(stage)::value;
    if (kt + 1 < nk)
      wait_vmcnt<2 * NJ>();
    else
      wait_vmcnt<0>();
    __builtin_amdgcn_s_barrier();
    if (kt + 2 < nk) issue(kt + 2, dma_stage_a<(S + 2) % 3>(sm), dma_stage_b<(S + 2) % 3>(sm));
    compute(dma_stage_a<S>(sm), dma_stage_b<S>(sm));
  };
  if (nk > 0) issue(0, sm.a0, sm.b0);
  if (nk > 1) issue(1, sm.a1, sm.b1);
  for (int kt = 0; kt < nk; kt += 3) {
    step(std::integral_constant<int, 0>{}, kt);
    if (kt + 1 < nk) step(std::integral_constant<int, 1>{}, kt + 1);
    if (kt + 2 < nk) step(std::integral_constant<int, 2>{}, kt + 2);
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int n = n0 + wn * WCOL + j * 16 + row16;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * 64 + i * 16 + MF::row(kq, r);
        TO* dst = C + (long)m * g.ldc + n;
        TO v{static_cast<TOR>(cr[i][j][r]), static_cast<TOR>(ci[i][j][r])};
        if (g.accumulate) {
          v.re += dst->re;
          v.im += dst->im;
        }
        *dst = v;
      }
    }
}

// Complex K-major GEMM: C[m][n] = sum_k A[k][m] * B'[k][n] with B' = B or conj(B), complex
// elements stored interleaved. Real and imaginary parts sit in separate LDS planes and each
// complex fragment product is four real MFMAs (re += ar br - ai bi, im += ar bi + ai br), so the
// complex flop count is 4x the real one at the same matrix-core rate. TA/TB/TO are storage types
// (cpx<double> or cpx<float>); TR is the compute type of the MFMA.
template <typename TR, typename TA, typename TB, typename TO, int BM, int BN, int BK, int WM, int WN>
__global__ void __launch_bounds__(WM* WN * 64)
    gemm_tn_cplx_kernel(GemmArgs g, int conj_b) {
  using MF = Mfma<TR>;
  using acc_t = typename MF::acc_t;
  constexpr int NT = WM * WN * 64;
  constexpr int PAD = 16;
  constexpr int LDA_S = BM + PAD;
  constexpr int LDB_S = BN + PAD;
  constexpr int FM = BM / WM / 16;
  constexpr int FN = BN / WN / 16;
  constexpr int EA = BK * BM / NT;
  constexpr int EB = BK * BN / NT;
  static_assert(BK % 4 == 0 && (BK * BM) % NT == 0 && (BK * BN) % NT == 0, "tile shape");
  __shared__ TR Ar[2][BK * LDA_S], Ai[2][BK * LDA_S];
  __shared__ TR Br[2][BK * LDB_S], Bi[2][BK * LDB_S];

  const int M = g.M, N = g.N, K = g.K;
  const int nbm = (M + BM - 1) / BM, nbn = (N + BN - 1) / BN;
  int tm, tn;
  tile_of(blockIdx.x, nbm, nbn, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const long bz = blockIdx.y;
  const TA* __restrict__ A = static_cast<const TA*>(g.A) + bz * g.strideA;
  const TB* __restrict__ B = static_cast<const TB*>(g.B) + bz * g.strideB;
  TO* __restrict__ C = static_cast<TO*>(g.C) + bz * g.strideC;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave / WN, wn = wave % WN, row16 = lane & 15, kq = lane >> 4;
  const TR bsign = conj_b ? TR(-1) : TR(1);

  acc_t cr[FM][FN], ci[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) cr[i][j] = ci[i][j] = acc_t{};

  TR rar[EA], rai[EA], rbr[EB], rbi[EB];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < EA; ++i) {
      const int e = t + i * NT, kr = e / BM, mc = e % BM, k = k0 + kr, m = m0 + mc;
      if (k < K && m < M) {
        const TA v = A[(long)k * g.lda + m];
        rar[i] = static_cast<TR>(v.re);
        rai[i] = static_cast<TR>(v.im);
      } else {
        rar[i] = rai[i] = TR(0);
      }
    }
#pragma unroll
    for (int i = 0; i < EB; ++i) {
      const int e = t + i * NT, kr = e / BN, nc = e % BN, k = k0 + kr, n = n0 + nc;
      if (k < K && n < N) {
        const TB v = B[(long)k * g.ldb + n];
        rbr[i] = static_cast<TR>(v.re);
        rbi[i] = bsign * static_cast<TR>(v.im);
      } else {
        rbr[i] = rbi[i] = TR(0);
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int i = 0; i < EA; ++i) {
      const int e = t + i * NT, o = (e / BM) * LDA_S + e % BM;
      Ar[buf][o] = rar[i];
      Ai[buf][o] = rai[i];
    }
#pragma unroll
    for (int i = 0; i < EB; ++i) {
      const int e = t + i * NT, o = (e / BN) * LDB_S + e % BN;
      Br[buf][o] = rbr[i];
      Bi[buf][o] = rbi[i];
    }
  };
  const int nk = (K + BK - 1) / BK;
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
#pragma unroll
    for (int kk = 0; kk < BK; kk += 4) {
      TR ar[FM], ai[FM], nai[FM], br[FN], bi[FN];
#pragma unroll
      for (int i = 0; i < FM; ++i) {
        const int o = (kk + kq) * LDA_S + wm * (BM / WM) + i * 16 + row16;
        ar[i] = Ar[cur][o];
        ai[i] = Ai[cur][o];
        nai[i] = -ai[i];
      }
#pragma unroll
      for (int j = 0; j < FN; ++j) {
        const int o = (kk + kq) * LDB_S + wn * (BN / WN) + j * 16 + row16;
        br[j] = Br[cur][o];
        bi[j] = Bi[cur][o];
      }
#pragma unroll
      for (int i = 0; i < FM; ++i)
#pragma unroll
        for (int j = 0; j < FN; ++j) {
          cr[i][j] = MF::run(ar[i], br[j], cr[i][j]);
          cr[i][j] = MF::run(nai[i], bi[j], cr[i][j]);
          ci[i][j] = MF::run(ar[i], bi[j], ci[i][j]);
          ci[i][j] = MF::run(ai[i], br[j], ci[i][j]);
        }
    }
    if (kt + 1 < nk) store(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int n = n0 + wn * (BN / WN) + j * 16 + row16;
      if (n >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * (BM / WM) + i * 16 + MF::row(kq, r);
        if (m < M) {
          TO* dst = C + (long)m * g.ldc + n;
          TO v(cr[i][j][r], ci[i][j][r]);
          if (g.accumulate) v = TO(dst->re + v.re, dst->im + v.im);
          *dst = v;
        }
      }
    }
}

// BF16 K-major GEMM on v_mfma_f32_16x16x32_bf16 (float accumulation): C = A^T B with A (K x M)
// and B (K x N) stored K-major in TA / TB (double, float or bf16) and rounded to bf16 on the way
// into LDS. For this MFMA lane l supplies 8 consecutive k of row (or column) l % 16 starting at
// k = 8 (l / 16), so the LDS tiles are held k-contiguous, As[m][k] and Bs[n][k] with rows of
// BK = 32 (+8 pad) values, and every fragment is one 16-byte LDS read. Result layout as the f32
// 16x16x4 MFMA: lane l, item r holds D[4 (l / 16) + r][l % 16].
template <typename TA, typename TB, typename TO, int BM, int BN, int WM, int WN, int VA, int VB>
__global__ void __launch_bounds__(WM* WN * 64)
    gemm_tn_bf16_kernel(GemmArgs g) {
  constexpr int BK = 32;
  constexpr int LDK = BK + 8;
  constexpr int NT = WM * WN * 64;
  constexpr int FM = BM / WM / 16;
  constexpr int FN = BN / WN / 16;
  constexpr int NA = BK * BM / (VA * NT);
  constexpr int NB = BK * BN / (VB * NT);
  static_assert(FM >= 1 && FN >= 1 && NA >= 1 && NB >= 1, "bf16 tile shape");
  static_assert((BK * BM) % (VA * NT) == 0 && (BK * BN) % (VB * NT) == 0, "bf16 tile loads");
  __shared__ __align__(16) __bf16 As[2][BM * LDK];
  __shared__ __align__(16) __bf16 Bs[2][BN * LDK];

  const int M = g.M, N = g.N, K = g.K;
  const int nbm = (M + BM - 1) / BM, nbn = (N + BN - 1) / BN;
  int tm, tn;
  tile_of(blockIdx.x, nbm, nbn, tm, tn);
  const int m0 = tm * BM, n0 = tn * BN;
  const long bz = blockIdx.y;
  const TA* __restrict__ A = static_cast<const TA*>(g.A) + bz * g.strideA;
  const TB* __restrict__ B = static_cast<const TB*>(g.B) + bz * g.strideB;
  TO* __restrict__ C = static_cast<TO*>(g.C) + bz * g.strideC;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave / WN, wn = wave % WN, row16 = lane & 15, kq = lane >> 4;

  f4 acc[FM][FN];
#pragma unroll
  for (int i = 0; i < FM; ++i)
#pragma unroll
    for (int j = 0; j < FN; ++j) acc[i][j] = f4{};

  __bf16 ra[NA * VA], rb[NB * VB];
  auto load = [&](int k0) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int e = t + i * NT, kr = e / (BM / VA), mc = (e % (BM / VA)) * VA;
      const int k = k0 + kr, m = m0 + mc;
      if constexpr (VA == 1) {
        ra[i] = (k < K && m < M) ? static_cast<__bf16>(static_cast<float>(A[(long)k * g.lda + m]))
                                 : static_cast<__bf16>(0.0f);
      } else {
        typedef TA vec_t __attribute__((ext_vector_type(VA)));
        if (k < K && m < M) {
          const vec_t x = *reinterpret_cast<const vec_t*>(A + (long)k * g.lda + m);
#pragma unroll
          for (int v = 0; v < VA; ++v) ra[i * VA + v] = static_cast<__bf16>(static_cast<float>(x[v]));
        } else {
#pragma unroll
          for (int v = 0; v < VA; ++v) ra[i * VA + v] = static_cast<__bf16>(0.0f);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int e = t + i * NT, kr = e / (BN / VB), nc = (e % (BN / VB)) * VB;
      const int k = k0 + kr, n = n0 + nc;
      if constexpr (VB == 1) {
        rb[i] = (k < K && n < N) ? static_cast<__bf16>(static_cast<float>(B[(long)k * g.ldb + n]))
                                 : static_cast<__bf16>(0.0f);
      } else {
        typedef TB vec_t __attribute__((ext_vector_type(VB)));
        if (k < K && n < N) {
          const vec_t x = *reinterpret_cast<const vec_t*>(B + (long)k * g.ldb + n);
#pragma unroll
          for (int v = 0; v < VB; ++v) rb[i * VB + v] = static_cast<__bf16>(static_cast<float>(x[v]));
        } else {
#pragma unroll
          for (int v = 0; v < VB; ++v) rb[i * VB + v] = static_cast<__bf16>(0.0f);
        }
      }
    }
  };
  auto store = [&](int buf) {
#pragma unroll
    for (int i = 0; i < NA; ++i) {
      const int e = t + i * NT, kr = e / (BM / VA), mc = (e % (BM / VA)) * VA;
#pragma unroll
      for (int v = 0; v < VA; ++v) As[buf][(mc + v) * LDK + kr] = ra[i * VA + v];
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
      const int e = t + i * NT, kr = e / (BN / VB), nc = (e % (BN / VB)) * VB;
#pragma unroll
      for (int v = 0; v < VB; ++v) Bs[buf][(nc + v) * LDK + kr] = rb[i * VB + v];
    }
  };
  const int nk = (K + BK - 1) / BK;
  if (nk > 0) {
    load(0);
    store(0);
  }
  __syncthreads();
  for (int kt = 0; kt < nk; ++kt) {
    const int cur = kt & 1;
    if (kt + 1 < nk) load((kt + 1) * BK);
    bf8 af[FM], bfr[FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
      af[i] = *reinterpret_cast<const bf8*>(&As[cur][(wm * (BM / WM) + i * 16 + row16) * LDK + kq * 8]);
#pragma unroll
    for (int j = 0; j < FN; ++j)
      bfr[j] = *reinterpret_cast<const bf8*>(&Bs[cur][(wn * (BN / WN) + j * 16 + row16) * LDK + kq * 8]);
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j)
        acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(af[i], bfr[j], acc[i][j], 0, 0, 0);
    if (kt + 1 < nk) store(cur ^ 1);
    __syncthreads();
  }
#pragma unroll
  for (int i = 0; i < FM; ++i) {
#pragma unroll
    for (int j = 0; j < FN; ++j) {
      const int n = n0 + wn * (BN / WN) + j * 16 + row16;
      if (n >= N) continue;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int m = m0 + wm * (BM / WM) + i * 16 + 4 * kq + r;
        if (m < M) {
          TO* dst = C + (long)m * g.ldc + n;
          if (g.accumulate)
            *dst = static_cast<TO>(static_cast<float>(*dst) + acc[i][j][r]);
          else
            *dst = static_cast<TO>(acc[i][j][r]);
        }
      }
    }
  }
}

// Host launcher of the bf16 GEMM: tile 1 = 128x128 (4 waves), 2 = 64x64 (4 waves); 0 = automatic.
template <typename TA, typename TB, typename TO>
int launch_gemm_bf16(const GemmArgs& g, int batch, hipStream_t stream, int tile_hint = 0);

template <typename TR, typename TA, typename TB, typename TO>
int launch_gemm_tn_cplx(const GemmArgs& g, int batch, int conj_b, hipStream_t stream,
                        int tile_hint = 0);

// Host-side launcher: picks a tile shape for the problem size and launches on `stream`.
// Returns a TNH_* status.
template <typename TC, typename TA, typename TB, typename TO>
int launch_gemm_tn(const GemmArgs& g, int batch, hipStream_t stream, int tile_hint = 0);

// Routes a GEMM to the real or the complex kernel by the compute type.
template <typename T>
struct real_of {
  using type = T;
  static constexpr bool cplx = false;
};
template <typename T>
struct real_of<cpx<T>> {
  using type = T;
  static constexpr bool cplx = true;
};

template <typename TC, typename TA, typename TB, typename TO>
inline int gemm_any(const GemmArgs& g, int batch, int conj_b, hipStream_t st, int tile) {
  if constexpr (real_of<TC>::cplx)
    return launch_gemm_tn_cplx<typename real_of<TC>::type, TA, TB, TO>(g, batch, conj_b, st, tile);
  else
    return launch_gemm_tn<TC, TA, TB, TO>(g, batch, st, tile);
}

}  // namespace tnh
