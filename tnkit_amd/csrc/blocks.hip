// Block-sparse (U(1)) contractions: the run gather that stacks an operand into per-charge GEMM
// matrices, and the persistent grouped GEMM that multiplies all charge groups of one
// contraction in a single launch.
//
// Layouts (decided on the host, see tnkit_amd/blocks.py):
//   gather : dst[j] = src[sbeg[r] + (j - dbeg[r]) * sstep[r]] for the run r covering j (runs
//            sorted by dbeg, disjoint); positions no run covers are written as zero, so padded
//            matrices come out with zero padding.
//   grouped GEMM, group g with descriptor (qa, lda, qb, ldb, po, R, K, C):
//            C[po + r*C + c] = sum_k A[qa + k*lda + r] * B[qb + k*ldb + c]     (K = 0: zeros)
//            A and B are the K-major stacks written by the gather (lda, ldb, qa, qb multiples of 4
//            elements, padding zero), so every tile row is read with 16-byte vector loads.
//   tiles  : the host lists (group, tile row, tile column) per position; workgroup b of a grid of
//            G (a multiple of 8) takes positions b, b + G, b + 2G, ... . Position p runs on XCD
//            p % 8, which the host uses to keep a panel of tiles that share an A row block on one
//            XCD's L2; entries with group < 0 are padding.
#include "gemm_tn.h"

namespace tnh {

namespace {

constexpr int kGatherThreads = 256;
constexpr int kGatherChunk = 2048;  // output elements per workgroup

// A workgroup owns output chunk [base, base + kGatherChunk); runs are sorted by destination and
// do not overlap, and clo/chi bound the runs that may touch the chunk. The descriptors of those
// runs are first copied into LDS (when there are at most kRunsLds of them, the usual case: runs
// shorter than a few dozen elements go to the tile kernel instead), then each thread takes 8
// elements 256 apart, finds each one's run by binary search in LDS, and issues all 8 loads
// before its first store. Searching in global memory instead puts three dependent global loads
// in front of every element. Chunks touched by more runs search in global memory.
constexpr int kRunsLds = 256;
constexpr int kPerThread = kGatherChunk / kGatherThreads;

template <typename TS, typename TD>
__global__ void __launch_bounds__(kGatherThreads)
    gather_runs_kernel(const TS* __restrict__ src, TD* __restrict__ dst, long n,
                       const long* __restrict__ sbeg, const long* __restrict__ sstep,
                       const long* __restrict__ dbeg, const long* __restrict__ len,
                       const int* __restrict__ clo, const int* __restrict__ chi, int fill) {
  __shared__ long s_d0[kRunsLds], s_d1[kRunsLds], s_s0[kRunsLds], s_st[kRunsLds];
  const long base = static_cast<long>(blockIdx.x) * kGatherChunk;
  const int lo = clo[blockIdx.x];
  const int hi = chi[blockIdx.x];  // runs [lo, hi) may overlap this chunk
  const int nr = hi - lo;
  const int t = threadIdx.x;
  TD v[kPerThread];
  bool hit[kPerThread];
  if (nr <= kRunsLds) {
    for (int k = t; k < nr; k += kGatherThreads) {
      const long d0 = dbeg[lo + k];
      s_d0[k] = d0;
      s_d1[k] = d0 + len[lo + k];
      s_s0[k] = sbeg[lo + k];
      s_st[k] = sstep[lo + k];
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
      const long j = base + t + static_cast<long>(i) * kGatherThreads;
      int a = 0, b = nr;  // last run with d0 <= j
      while (b - a > 1) {
        const int mid = (a + b) >> 1;
        if (s_d0[mid] <= j)
          a = mid;
        else
          b = mid;
      }
      hit[i] = j < n && a < nr && s_d0[a] <= j && j < s_d1[a];
      v[i] = hit[i] ? static_cast<TD>(src[s_s0[a] + (j - s_d0[a]) * s_st[a]]) : TD(0);
    }
  } else {
    int r = lo;
#pragma unroll
    for (int i = 0; i < kPerThread; ++i) {
      const long j = base + t + static_cast<long>(i) * kGatherThreads;
      int a = r, b = hi;  // j grows with i, so r only moves forward
      while (b - a > 1) {
        const int mid = (a + b) >> 1;
        if (dbeg[mid] <= j)
          a = mid;
        else
          b = mid;
      }
      r = a;
      const long d0 = a < hi ? dbeg[a] : 0;
      hit[i] = j < n && a < hi && d0 <= j && j < d0 + len[a];
      v[i] = hit[i] ? static_cast<TD>(src[sbeg[a] + (j - d0) * sstep[a]]) : TD(0);
    }
  }
#pragma unroll
  for (int i = 0; i < kPerThread; ++i) {
    const long j = base + t + static_cast<long>(i) * kGatherThreads;
    if (j < n && (hit[i] || fill)) dst[j] = v[i];
  }
}

template <typename TS, typename TD>
int launch_gather(long n, const void* src, void* dst, long nruns, const long* runs,
                  const int* chunks, int fill, hipStream_t stream) {
  if (n == 0) return TNH_OK;
  const long nchunks = (n + kGatherChunk - 1) / kGatherChunk;
  TNH_REQUIRE(nchunks < (1L << 31), TNH_ERR_ARG, "gather_runs: too many elements");
  hipLaunchKernelGGL((gather_runs_kernel<TS, TD>), dim3(static_cast<unsigned>(nchunks)),
                     dim3(kGatherThreads), 0, stream, static_cast<const TS*>(src),
                     static_cast<TD*>(dst), n, runs, runs + nruns, runs + 2 * nruns,
                     runs + 3 * nruns, chunks, chunks + nchunks, fill);
  TNH_CHECK_LAUNCH("gather_runs");
  return TNH_OK;
}

// Tiled gather: tile t = (src0, srow, scol, dst0, drow, nrows, ncols) moves
//   dst[dst0 + i*drow + j] = src[src0 + i*srow + j*scol],   i < nrows <= 64, j < ncols <= 64.
// With scol == 1 (rows contiguous on both sides: many short runs moved together) the tile is
// copied directly, its elements spread over all lanes; otherwise it is a transpose staged through
// a 64 x 65 LDS tile, read along i (contiguous in the source when srow == 1) and written along j
// (contiguous in the destination).
constexpr int kTile = 64;

template <typename TS, typename TD>
__global__ void __launch_bounds__(256)
    gather_tiles_kernel(const TS* __restrict__ src, TD* __restrict__ dst,
                        const long* __restrict__ tiles) {
  __shared__ TD buf[kTile][kTile + 1];
  const long* tl = tiles + 7 * static_cast<long>(blockIdx.x);
  const long src0 = tl[0], srow = tl[1], scol = tl[2], dst0 = tl[3], drow = tl[4];
  const int nrows = static_cast<int>(tl[5]), ncols = static_cast<int>(tl[6]);
  const int t = threadIdx.x;
  // fixed trip counts (a tile has at most 64 x 64 = 16 x 256 elements), fully unrolled, so a
  // thread's 16 loads are all issued before the first store
  constexpr int kPer = kTile * kTile / 256;
  if (scol == 1) {
    const int n = nrows * ncols;
    TD v[kPer];
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int e = t + q * 256;
      const int i = e / ncols, j = e - i * ncols;
      v[q] = e < n ? static_cast<TD>(src[src0 + static_cast<long>(i) * srow + j]) : TD(0);
    }
#pragma unroll
    for (int q = 0; q < kPer; ++q) {
      const int e = t + q * 256;
      const int i = e / ncols, j = e - i * ncols;
      if (e < n) dst[dst0 + static_cast<long>(i) * drow + j] = v[q];
    }
    return;
  }
  const int i = t & 63;
  const TS* s = src + src0 + static_cast<long>(i) * srow;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int j = (t >> 6) + 4 * q;
    if (i < nrows && j < ncols) buf[j][i] = static_cast<TD>(s[static_cast<long>(j) * scol]);
  }
  __syncthreads();
  const int jj = t & 63;
  TD* d = dst + dst0 + jj;
#pragma unroll
  for (int q = 0; q < kPer; ++q) {
    const int ii = (t >> 6) + 4 * q;
    if (ii < nrows && jj < ncols) d[static_cast<long>(ii) * drow] = buf[jj][ii];
  }
}

template <typename TS, typename TD>
int launch_tiles(long ntiles, const void* src, void* dst, const long* tiles, hipStream_t stream) {
  if (ntiles == 0) return TNH_OK;
  TNH_REQUIRE(ntiles < (1L << 31), TNH_ERR_ARG, "gather_tiles: too many tiles");
  hipLaunchKernelGGL((gather_tiles_kernel<TS, TD>), dim3(static_cast<unsigned>(ntiles)), dim3(256),
                     0, stream, static_cast<const TS*>(src), static_cast<TD*>(dst), tiles);
  TNH_CHECK_LAUNCH("gather_tiles");
  return TNH_OK;
}

// Small grouped products (every group R <= 8 and K <= 8, the MPO contractions of the two-site
// matvec): one thread per output column, the group's A (K x R) in registers, B rows read and
// C rows written contiguously across the wavefront; memory bound, no matrix cores.
// chunks (device, nchunks x 2 ints): (group, first column) of each workgroup's 1024 columns.
constexpr int kSmall = 8;
constexpr int kSmallCols = 1024;

template <typename T>
__global__ void __launch_bounds__(256)
    grouped_small_kernel(const long* __restrict__ desc, const int* __restrict__ chunks,
                         const T* __restrict__ A, const T* __restrict__ B, T* __restrict__ C) {
  const int g = chunks[2 * blockIdx.x];
  const int c0 = chunks[2 * blockIdx.x + 1];
  const long* d = desc + 8 * static_cast<long>(g);
  const T* Ag = A + d[0];
  const long lda = d[1];
  const T* Bg = B + d[2];
  const long ldb = d[3];
  T* Cg = C + d[4];
  const int R = static_cast<int>(d[5]), K = static_cast<int>(d[6]), N = static_cast<int>(d[7]);
  T a[kSmall][kSmall];
#pragma unroll
  for (int k = 0; k < kSmall; ++k)
#pragma unroll
    for (int r = 0; r < kSmall; ++r) a[k][r] = (k < K && r < R) ? Ag[k * lda + r] : T(0);
  for (int q = 0; q < kSmallCols / 256; ++q) {
    const int c = c0 + q * 256 + threadIdx.x;
    if (c >= N || c >= c0 + kSmallCols) break;
    T b[kSmall];
#pragma unroll
    for (int k = 0; k < kSmall; ++k) b[k] = k < K ? Bg[k * ldb + c] : T(0);
#pragma unroll
    for (int r = 0; r < kSmall; ++r) {
      if (r < R) {
        T acc = T(0);
#pragma unroll
        for (int k = 0; k < kSmall; ++k) acc += a[k][r] * b[k];
        Cg[static_cast<long>(r) * N + c] = acc;
      }
    }
  }
}

// The same product for groups with R <= 32 and K <= 32 (the two MPO tensors of the two-site
// matvec contracted into one six-leg operator, whose (w, s1, s2) x (s1', s2', w') charge groups
// have up to D d^2 rows and columns): the group's A (K x R) is staged in LDS and read as a
// broadcast, each thread keeps one column of B (K values) in registers and writes its R outputs;
// B rows are read and C rows written contiguously across the wavefront.
constexpr int kMid = 32;

template <typename T>
__global__ void __launch_bounds__(256)
    grouped_mid_kernel(const long* __restrict__ desc, const int* __restrict__ chunks,
                       const T* __restrict__ A, const T* __restrict__ B, T* __restrict__ C) {
  __shared__ T a[kMid][kMid];
  const int g = chunks[2 * blockIdx.x];
  const int c0 = chunks[2 * blockIdx.x + 1];
  const long* d = desc + 8 * static_cast<long>(g);
  const T* Ag = A + d[0];
  const long lda = d[1];
  const T* Bg = B + d[2];
  const long ldb = d[3];
  T* Cg = C + d[4];
  const int R = static_cast<int>(d[5]), K = static_cast<int>(d[6]), N = static_cast<int>(d[7]);
  for (int e = threadIdx.x; e < kMid * kMid; e += 256) {
    const int k = e / kMid, r = e % kMid;
    a[k][r] = (k < K && r < R) ? Ag[k * lda + r] : T(0);
  }
  __syncthreads();
  for (int q = 0; q < kSmallCols / 256; ++q) {
    const int c = c0 + q * 256 + threadIdx.x;
    if (c >= N || c >= c0 + kSmallCols) break;
    T b[kMid];
#pragma unroll
    for (int k = 0; k < kMid; ++k) b[k] = k < K ? Bg[k * ldb + c] : T(0);
    for (int r = 0; r < R; ++r) {
      T acc = T(0);
#pragma unroll
      for (int k = 0; k < kMid; ++k) acc += a[k][r] * b[k];
      Cg[static_cast<long>(r) * N + c] = acc;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// grouped GEMM
// ---------------------------------------------------------------------------------------------

struct GroupArgs {
  const long* desc;  // ngroups x 8: qa, lda, qb, ldb, po, R, K, C
  const int* tiles;  // ntiles x 3: group, tile row, tile column
  long ntiles;
  const void* A;
  const void* B;
  void* C;
};

template <typename T, int BM, int BN, int BK, int WM, int WN>
__global__ void __launch_bounds__(WM* WN * 64) grouped_gemm_kernel(GroupArgs ga) {
  using MF = Mfma<T>;
  using acc_t = typename MF::acc_t;
  constexpr int NT = WM * WN * 64;
  constexpr int PAD = 16;
  constexpr int LDA_S = BM + PAD;
  constexpr int LDB_S = BN + PAD;
  constexpr int FM = BM / WM / 16;
  constexpr int FN = BN / WN / 16;
  constexpr int V = 16 / sizeof(T);
  // 16-byte loads where the tile splits evenly into them (the skinny 16-row A tile does not)
  constexpr int VA = (BK * BM) % (V * NT) == 0 ? V : 1;
  constexpr int VB = (BK * BN) % (V * NT) == 0 ? V : 1;
  static_assert(BK % 4 == 0 && FM >= 1 && FN >= 1, "tile shape");

  __shared__ T As[2][BK * LDA_S];
  __shared__ T Bs[2][BK * LDB_S];

  const int t = threadIdx.x;
  const int lane = t & 63;
  const int wave = t >> 6;
  const int wm = wave / WN;
  const int wn = wave % WN;
  const int row16 = lane & 15;
  const int kq = lane >> 4;

  for (long pos = blockIdx.x; pos < ga.ntiles; pos += gridDim.x) {
    const int g = ga.tiles[3 * pos];
    if (g < 0) continue;
    const int m0 = ga.tiles[3 * pos + 1] * BM;
    const int n0 = ga.tiles[3 * pos + 2] * BN;
    const long* d = ga.desc + 8 * static_cast<long>(g);
    const T* __restrict__ A = static_cast<const T*>(ga.A) + d[0];
    const long lda = d[1];
    const T* __restrict__ B = static_cast<const T*>(ga.B) + d[2];
    const long ldb = d[3];
    T* __restrict__ C = static_cast<T*>(ga.C) + d[4];
    const int M = static_cast<int>(d[5]);
    const int K = static_cast<int>(d[6]);
    const int N = static_cast<int>(d[7]);

    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};

    // vector loads may run past M (or N) into the zero padding of the row, never past lda/ldb
    TileLoader<T, T, BK, BM, NT, VA> la;
    TileLoader<T, T, BK, BN, NT, VB> lb;
    const int nk = (K + BK - 1) / BK;
    if (nk > 0) {
      la.load(A, lda, 0, K, m0, M, t);
      lb.load(B, ldb, 0, K, n0, N, t);
      la.store(As[0], LDA_S, t);
      lb.store(Bs[0], LDB_S, t);
    }
    __syncthreads();
    for (int kt = 0; kt < nk; ++kt) {
      const int cur = kt & 1;
      if (kt + 1 < nk) {
        la.load(A, lda, (kt + 1) * BK, K, m0, M, t);
        lb.load(B, ldb, (kt + 1) * BK, K, n0, N, t);
      }
      const T* as = As[cur];
      const T* bs = Bs[cur];
#pragma unroll
      for (int kk = 0; kk < BK; kk += 4) {
        T af[FM], bf[FN];
#pragma unroll
        for (int i = 0; i < FM; ++i)
          af[i] = as[(kk + kq) * LDA_S + wm * (BM / WM) + i * 16 + row16];
#pragma unroll
        for (int j = 0; j < FN; ++j)
          bf[j] = bs[(kk + kq) * LDB_S + wn * (BN / WN) + j * 16 + row16];
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
          if (m < M) C[static_cast<long>(m) * N + n] = acc[i][j][r];
        }
      }
    }
  }
}

// The grouped GEMM with direct-to-LDS loads: the ring of gemm_dma_kernel (tn's three LDS stages,
// tiles moved by global_load_lds_dwordx4 with no VGPR staging, rows rotated by 16 elements per
// (k row mod 4) for distinct banks) on the persistent tile loop above. Requirements the host
// plans meet: every group's K-major stack has its K rows padded with zero rows to a multiple of
// 32 (so whole BK slabs are read), lda / ldb are multiples of 4 elements and the stacks 16-byte
// aligned. A chunk whose columns start past the padded row (a tile hanging over R or C) is
// loaded from the row's last chunk instead: it only feeds output rows or columns past R / C,
// which are not stored, and every load stays inside the group's rows.
template <typename T, int BM, int BN, int BK>
struct GroupDmaStages {
  T a0[BK * BM], a1[BK * BM], a2[BK * BM];
  T b0[BK * BN], b1[BK * BN], b2[BK * BN];
};

template <typename T, int BM, int BN, int BK, int WM, int WN>
__global__ void __launch_bounds__(WM* WN * 64) grouped_gemm_dma_kernel(GroupArgs ga) {
  using MF = Mfma<T>;
  using acc_t = typename MF::acc_t;
  constexpr int NT = WM * WN * 64;
  constexpr int FM = BM / WM / 16, FN = BN / WN / 16;
  constexpr int EPC = 16 / sizeof(T);             // elements per 16-byte chunk
  constexpr int CPA = BM / EPC, CPB = BN / EPC;   // chunks per k row of each tile
  constexpr int NJA = BK * CPA / NT, NJB = BK * CPB / NT;
  constexpr int ROT = 16 / EPC;                   // the 16-element rotation, in chunks
  static_assert(NJA >= 1 && NJB >= 1 && NJA * NT == BK * CPA && NJB * NT == BK * CPB &&
                    BK % 4 == 0 && FM >= 1 && FN >= 1,
                "grouped dma tile shape");
  __shared__ __align__(16) GroupDmaStages<T, BM, BN, BK> sm;
  const int t = threadIdx.x, lane = t & 63, wave = t >> 6;
  const int wm = wave / WN, wn = wave % WN, row16 = lane & 15, kq = lane >> 4;
  const int aoff = wm * (BM / WM) + row16 + 16 * kq, boff = wn * (BN / WN) + row16 + 16 * kq;

  for (long pos = blockIdx.x; pos < ga.ntiles; pos += gridDim.x) {
    const int g = ga.tiles[3 * pos];
    if (g < 0) continue;
    const int m0 = ga.tiles[3 * pos + 1] * BM;
    const int n0 = ga.tiles[3 * pos + 2] * BN;
    const long* d = ga.desc + 8 * static_cast<long>(g);
    const T* __restrict__ A = static_cast<const T*>(ga.A) + d[0];
    const long lda = d[1];
    const T* __restrict__ B = static_cast<const T*>(ga.B) + d[2];
    const long ldb = d[3];
    T* __restrict__ C = static_cast<T*>(ga.C) + d[4];
    const int M = static_cast<int>(d[5]);
    const int K = static_cast<int>(d[6]);
    const int N = static_cast<int>(d[7]);
    long goffA[NJA], goffB[NJB];
#pragma unroll
    for (int j = 0; j < NJA; ++j) {
      const int c = j * NT + t, r = c / CPA, p = c % CPA;
      const int q = ((p - ROT * (r & 3)) % CPA + CPA) % CPA;
      const long col = min(static_cast<long>(m0 + EPC * q), lda - EPC);
      goffA[j] = static_cast<long>(r) * lda + col;
    }
#pragma unroll
    for (int j = 0; j < NJB; ++j) {
      const int c = j * NT + t, r = c / CPB, p = c % CPB;
      const int q = ((p - ROT * (r & 3)) % CPB + CPB) % CPB;
      const long col = min(static_cast<long>(n0 + EPC * q), ldb - EPC);
      goffB[j] = static_cast<long>(r) * ldb + col;
    }
    acc_t acc[FM][FN];
#pragma unroll
    for (int i = 0; i < FM; ++i)
#pragma unroll
      for (int j = 0; j < FN; ++j) acc[i][j] = acc_t{};
    const int nk = (K + BK - 1) / BK;
    auto issue = [&](int kt, T* as, T* bs) __attribute__((always_inline)) {
      const T* a = A + static_cast<long>(kt) * BK * lda;
      const T* b = B + static_cast<long>(kt) * BK * ldb;
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
        for (int j = 0; j < FN; ++j) bf[j] = bs[(kk + kq) * BN + ((boff + j * 16) & (BN - 1))];
#pragma unroll
        for (int i = 0; i < FM; ++i)
#pragma unroll
          for (int j = 0; j < FN; ++j) acc[i][j] = MF::run(af[i], bf[j], acc[i][j]);
      }
    };
    auto step = [&](auto stage, int kt) __attribute__((always_inline)) {
      constexpr int S = decltype(stage)::value;
      if (kt + 1 < nk)
        wait_vmcnt<NJA + NJB>();
      else
        wait_vmcnt<0>();
      __builtin_amdgcn_s_barrier();
      if (kt + 2 < nk) issue(kt + 2, dma_stage_a<(S + 2) % 3>(sm), dma_stage_b<(S + 2) % 3>(sm));
      compute(dma_stage_a<S>(sm), dma_stage_b<S>(sm));
    };
    // the previous tile's stores and every wave's reads of the stages are done before the
    // stages are refilled
    wait_vmcnt<0>();
    __syncthreads();
    if (nk > 0) issue(0, sm.a0, sm.b0);
    if (nk > 1) issue(1, sm.a1, sm.b1);
    for (int kt = 0; kt < nk; kt += 3) {
      step(std::integral_constant<int, 0>{}, kt);
      if (kt + 1 < nk) step(std::integral_constant<int, 1>{}, kt + 1);
      if (kt + 2 < nk) step(std::integral_constant<int, 2>{}, kt + 2);
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
          if (m < M) C[static_cast<long>(m) * N + n] = acc[i][j][r];
        }
      }
    }
  }
}

struct TileShape {
  int bm, bn, bk, wm, wn;
};

template <typename T>
TileShape tile_shape(int cfg) {
  constexpr bool dbl = sizeof(T) == 8;
  switch (cfg) {
    case 1: return {128, 128, dbl ? 16 : 32, 2, 2};
    case 2: return {64, 64, 16, 2, 2};
    case 3: return {32, 32, 16, 2, 2};
    case 4: return {16, 128, 16, 1, 4};
    case 5: return {64, 128, 16, 2, 2};
    case 11: return {128, 128, dbl ? 16 : 32, 2, 2};
    case 12: return {64, 64, dbl ? 16 : 32, 2, 2};
    case 13: return {32, 32, dbl ? 16 : 32, 2, 2};
    case 15: return {64, 128, dbl ? 16 : 32, 2, 2};
    default: return {0, 0, 0, 0, 0};
  }
}

template <typename T, int BM, int BN, int BK, int WM, int WN>
int launch_grouped_cfg(const GroupArgs& ga, int grid, hipStream_t stream, int* max_grid) {
  auto kern = grouped_gemm_kernel<T, BM, BN, BK, WM, WN>;
  if (max_grid) {  // the persistent grid: resident workgroups on every CU, a multiple of 8
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, WM * WN * 64, 0) !=
            hipSuccess) {
      set_error("grouped_gemm: device query failed");
      return TNH_ERR_HIP;
    }
    *max_grid = (cus * (per_cu > 0 ? per_cu : 1) + 7) / 8 * 8;
    return TNH_OK;
  }
  if (ga.ntiles == 0) return TNH_OK;
  hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(grid)), dim3(WM * WN * 64), 0, stream, ga);
  TNH_CHECK_LAUNCH("grouped_gemm");
  return TNH_OK;
}

template <typename T, int BM, int BN, int BK, int WM, int WN>
int launch_grouped_dma_cfg(const GroupArgs& ga, int grid, hipStream_t stream, int* max_grid) {
  auto kern = grouped_gemm_dma_kernel<T, BM, BN, BK, WM, WN>;
  if (max_grid) {
    int dev = 0, cus = 0, per_cu = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&per_cu, kern, WM * WN * 64, 0) !=
            hipSuccess) {
      set_error("grouped_gemm: device query failed");
      return TNH_ERR_HIP;
    }
    *max_grid = (cus * (per_cu > 0 ? per_cu : 1) + 7) / 8 * 8;
    return TNH_OK;
  }
  if (ga.ntiles == 0) return TNH_OK;
  hipLaunchKernelGGL(kern, dim3(static_cast<unsigned>(grid)), dim3(WM * WN * 64), 0, stream, ga);
  TNH_CHECK_LAUNCH("grouped_gemm_dma");
  return TNH_OK;
}

template <typename T>
int launch_grouped(int cfg, const GroupArgs& ga, int grid, hipStream_t stream, int* max_grid) {
  constexpr bool dbl = sizeof(T) == 8;
  constexpr int BKb = dbl ? 16 : 32;
  switch (cfg) {
    case 11: return launch_grouped_dma_cfg<T, 128, 128, BKb, 2, 2>(ga, grid, stream, max_grid);
    case 12: return launch_grouped_dma_cfg<T, 64, 64, BKb, 2, 2>(ga, grid, stream, max_grid);
    case 13: return launch_grouped_dma_cfg<T, 32, 32, BKb, 2, 2>(ga, grid, stream, max_grid);
    case 15: return launch_grouped_dma_cfg<T, 64, 128, BKb, 2, 2>(ga, grid, stream, max_grid);
    case 1: return launch_grouped_cfg<T, 128, 128, BKb, 2, 2>(ga, grid, stream, max_grid);
    case 2: return launch_grouped_cfg<T, 64, 64, 16, 2, 2>(ga, grid, stream, max_grid);
    case 3: return launch_grouped_cfg<T, 32, 32, 16, 2, 2>(ga, grid, stream, max_grid);
    case 4: return launch_grouped_cfg<T, 16, 128, 16, 1, 4>(ga, grid, stream, max_grid);
    case 5: return launch_grouped_cfg<T, 64, 128, 16, 2, 2>(ga, grid, stream, max_grid);
    default: set_error("grouped_gemm: unknown tile config %d", cfg); return TNH_ERR_ARG;
  }
}

}  // namespace
}  // namespace tnh

using namespace tnh;

extern "C" int tnh_gather_chunk(void) { return kGatherChunk; }

extern "C" int tnh_gather_runs(int src_dtype, int dst_dtype, long n, const void* src, void* dst,
                               long nruns, const long* runs, const int* chunks, int fill,
                               tnh_stream_t stream) {
  TNH_REQUIRE(n >= 0 && nruns >= 0, TNH_ERR_ARG, "tnh_gather_runs: negative size");
  TNH_REQUIRE(n == 0 || (dst && chunks && (nruns == 0 || (src && runs))), TNH_ERR_ARG,
              "tnh_gather_runs: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (src_dtype == TNH_F64 && dst_dtype == TNH_F64)
    return launch_gather<double, double>(n, src, dst, nruns, runs, chunks, fill, s);
  if (src_dtype == TNH_F32 && dst_dtype == TNH_F32)
    return launch_gather<float, float>(n, src, dst, nruns, runs, chunks, fill, s);
  if (src_dtype == TNH_F64 && dst_dtype == TNH_F32)
    return launch_gather<double, float>(n, src, dst, nruns, runs, chunks, fill, s);
  if (src_dtype == TNH_F32 && dst_dtype == TNH_F64)
    return launch_gather<float, double>(n, src, dst, nruns, runs, chunks, fill, s);
  set_error("tnh_gather_runs: unsupported dtypes %d -> %d", src_dtype, dst_dtype);
  return TNH_ERR_UNSUPPORTED;
}

extern "C" int tnh_gather_tiles(int src_dtype, int dst_dtype, long ntiles, const void* src,
                                void* dst, const long* tiles, tnh_stream_t stream) {
  TNH_REQUIRE(ntiles >= 0, TNH_ERR_ARG, "tnh_gather_tiles: negative size");
  if (ntiles == 0) return TNH_OK;
  TNH_REQUIRE(src && dst && tiles, TNH_ERR_ARG, "tnh_gather_tiles: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (src_dtype == TNH_F64 && dst_dtype == TNH_F64)
    return launch_tiles<double, double>(ntiles, src, dst, tiles, s);
  if (src_dtype == TNH_F32 && dst_dtype == TNH_F32)
    return launch_tiles<float, float>(ntiles, src, dst, tiles, s);
  if (src_dtype == TNH_F64 && dst_dtype == TNH_F32)
    return launch_tiles<double, float>(ntiles, src, dst, tiles, s);
  if (src_dtype == TNH_F32 && dst_dtype == TNH_F64)
    return launch_tiles<float, double>(ntiles, src, dst, tiles, s);
  set_error("tnh_gather_tiles: unsupported dtypes %d -> %d", src_dtype, dst_dtype);
  return TNH_ERR_UNSUPPORTED;
}

extern "C" int tnh_grouped_small(int dtype, const long* desc, long nchunks, const int* chunks,
                                 const void* A, const void* B, void* C, tnh_stream_t stream) {
  TNH_REQUIRE(nchunks >= 0 && nchunks < (1L << 31), TNH_ERR_ARG,
              "tnh_grouped_small: bad chunk count");
  if (nchunks == 0) return TNH_OK;
  TNH_REQUIRE(desc && chunks && A && B && C, TNH_ERR_ARG, "tnh_grouped_small: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 grid(static_cast<unsigned>(nchunks));
  if (dtype == TNH_F64) {
    hipLaunchKernelGGL(grouped_small_kernel<double>, grid, dim3(256), 0, s, desc, chunks,
                       static_cast<const double*>(A), static_cast<const double*>(B),
                       static_cast<double*>(C));
  } else if (dtype == TNH_F32) {
    hipLaunchKernelGGL(grouped_small_kernel<float>, grid, dim3(256), 0, s, desc, chunks,
                       static_cast<const float*>(A), static_cast<const float*>(B),
                       static_cast<float*>(C));
  } else {
    set_error("tnh_grouped_small: unsupported dtype %d", dtype);
    return TNH_ERR_UNSUPPORTED;
  }
  TNH_CHECK_LAUNCH("grouped_small");
  return TNH_OK;
}

extern "C" int tnh_grouped_mid(int dtype, const long* desc, long nchunks, const int* chunks,
                               const void* A, const void* B, void* C, tnh_stream_t stream) {
  TNH_REQUIRE(nchunks >= 0 && nchunks < (1L << 31), TNH_ERR_ARG,
              "tnh_grouped_mid: bad chunk count");
  if (nchunks == 0) return TNH_OK;
  TNH_REQUIRE(desc && chunks && A && B && C, TNH_ERR_ARG, "tnh_grouped_mid: null pointer");
  hipStream_t s = static_cast<hipStream_t>(stream);
  const dim3 grid(static_cast<unsigned>(nchunks));
  if (dtype == TNH_F64) {
    hipLaunchKernelGGL(grouped_mid_kernel<double>, grid, dim3(256), 0, s, desc, chunks,
                       static_cast<const double*>(A), static_cast<const double*>(B),
                       static_cast<double*>(C));
  } else if (dtype == TNH_F32) {
    hipLaunchKernelGGL(grouped_mid_kernel<float>, grid, dim3(256), 0, s, desc, chunks,
                       static_cast<const float*>(A), static_cast<const float*>(B),
                       static_cast<float*>(C));
  } else {
    set_error("tnh_grouped_mid: unsupported dtype %d", dtype);
    return TNH_ERR_UNSUPPORTED;
  }
  TNH_CHECK_LAUNCH("grouped_mid");
  return TNH_OK;
}

extern "C" int tnh_grouped_tile(int dtype, int cfg, int* bm, int* bn, int* max_grid) {
  TileShape ts{};
  int rc = TNH_OK;
  if (dtype == TNH_F64) {
    ts = tile_shape<double>(cfg);
    if (ts.bm && max_grid) rc = launch_grouped<double>(cfg, GroupArgs{}, 0, nullptr, max_grid);
  } else if (dtype == TNH_F32) {
    ts = tile_shape<float>(cfg);
    if (ts.bm && max_grid) rc = launch_grouped<float>(cfg, GroupArgs{}, 0, nullptr, max_grid);
  } else {
    set_error("tnh_grouped_tile: unsupported dtype %d", dtype);
    return TNH_ERR_UNSUPPORTED;
  }
  TNH_REQUIRE(ts.bm > 0, TNH_ERR_ARG, "tnh_grouped_tile: unknown tile config %d", cfg);
  if (bm) *bm = ts.bm;
  if (bn) *bn = ts.bn;
  return rc;
}

extern "C" int tnh_grouped_gemm(int dtype, const long* desc, long ntiles, const int* tiles,
                                const void* A, const void* B, void* C, int tile_cfg, int grid,
                                tnh_stream_t stream) {
  TNH_REQUIRE(ntiles >= 0 && grid >= 0, TNH_ERR_ARG, "tnh_grouped_gemm: negative size");
  if (ntiles == 0) return TNH_OK;
  TNH_REQUIRE(grid > 0 && grid % 8 == 0, TNH_ERR_ARG,
              "tnh_grouped_gemm: grid %d must be a positive multiple of 8", grid);
  TNH_REQUIRE(desc && tiles && A && B && C, TNH_ERR_ARG, "tnh_grouped_gemm: null pointer");
  TNH_REQUIRE(reinterpret_cast<uintptr_t>(A) % 16 == 0 && reinterpret_cast<uintptr_t>(B) % 16 == 0,
              TNH_ERR_ARG, "tnh_grouped_gemm: A and B must be 16-byte aligned");
  GroupArgs ga{desc, tiles, ntiles, A, B, C};
  hipStream_t s = static_cast<hipStream_t>(stream);
  if (dtype == TNH_F64) return launch_grouped<double>(tile_cfg, ga, grid, s, nullptr);
  if (dtype == TNH_F32) return launch_grouped<float>(tile_cfg, ga, grid, s, nullptr);
  set_error("tnh_grouped_gemm: unsupported dtype %d", dtype);
  return TNH_ERR_UNSUPPORTED;
}
