// Tile selection and launch of the K-major MFMA GEMM, plus the C-ABI entry tnh_gemm_tn.
#include <cstdint>
#include <type_traits>

#include "gemm_tn.h"

namespace tnh {

namespace {

inline long ceil_div(long a, long b) { return (a + b - 1) / b; }

template <int V, typename T>
bool vec_ok(const void* p, long ld, long cols) {
  return V > 1 && ld % V == 0 && cols % V == 0 && (reinterpret_cast<uintptr_t>(p) % (V * sizeof(T))) == 0;
}

template <typename TC, typename TA, typename TB, typename TO, int BM, int BN, int BK, int WM, int WN,
          int VA, int VB>
int launch_v(const GemmArgs& g, int batch, hipStream_t stream) {
  const long tiles = ceil_div(g.M, BM) * ceil_div(g.N, BN);
  dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(batch));
  hipLaunchKernelGGL((gemm_tn_kernel<TC, TA, TB, TO, BM, BN, BK, WM, WN, VA, VB>), grid,
                     dim3(WM * WN * 64), 0, stream, g);
  TNH_CHECK_LAUNCH("gemm_tn");
  return TNH_OK;
}

// Picks 16-byte vector loads for an operand when its leading dimension, column count, base
// pointer and batch stride allow them.
template <typename TC, typename TA, typename TB, typename TO, int BM, int BN, int BK, int WM, int WN>
int launch_one(const GemmArgs& g, int batch, hipStream_t stream) {
  const long tiles = ceil_div(g.M, BM) * ceil_div(g.N, BN);
  if (tiles == 0 || batch == 0) return TNH_OK;
  TNH_REQUIRE(tiles < (1L << 31) && batch <= 65535, TNH_ERR_ARG, "gemm_tn: grid too large");
  constexpr int VA = 16 / sizeof(TA), VB = 16 / sizeof(TB);
  constexpr bool can_a = (BK * BM) % (VA * WM * WN * 64) == 0;
  constexpr bool can_b = (BK * BN) % (VB * WM * WN * 64) == 0;
  const bool va = can_a && vec_ok<VA, TA>(g.A, g.lda, g.M) && (batch == 1 || g.strideA % VA == 0);
  const bool vb = can_b && vec_ok<VB, TB>(g.B, g.ldb, g.N) && (batch == 1 || g.strideB % VB == 0);
  if constexpr (can_a && can_b) {
    if (va && vb) return launch_v<TC, TA, TB, TO, BM, BN, BK, WM, WN, VA, VB>(g, batch, stream);
  }
  if constexpr (can_a) {
    if (va) return launch_v<TC, TA, TB, TO, BM, BN, BK, WM, WN, VA, 1>(g, batch, stream);
  }
  if constexpr (can_b) {
    if (vb) return launch_v<TC, TA, TB, TO, BM, BN, BK, WM, WN, 1, VB>(g, batch, stream);
  }
  return launch_v<TC, TA, TB, TO, BM, BN, BK, WM, WN, 1, 1>(g, batch, stream);
}

// The direct-to-LDS kernel (gemm_dma_kernel) when the operands are stored in the compute type
// and the shape and strides fit it: float64 always (0.74 vs 0.79 ms on the chi = 1024 matvec
// GEMMs), float32 when there are at most two 128 x 128 tiles per CU (a long K, as the third
// matvec GEMM: 0.36 vs 0.43 ms); with more tiles the staged kernel keeps up and leads slightly.
template <typename TC, typename TA, typename TB, typename TO, int WNC = 2>
bool try_dma(const GemmArgs& g, int batch, hipStream_t stream, int& rc) {
  if constexpr (std::is_same<TA, TC>::value && std::is_same<TB, TC>::value &&
                (std::is_same<TC, float>::value || std::is_same<TC, double>::value)) {
    constexpr int EPC = 16 / sizeof(TC);
    constexpr int BK = std::is_same<TC, float>::value ? 32 : 16;
    if (batch < 1 || g.M % 128 || g.N % 128 || g.K < BK || g.K % BK) return false;
    if (g.lda % EPC || g.ldb % EPC || reinterpret_cast<uintptr_t>(g.A) % 16 ||
        reinterpret_cast<uintptr_t>(g.B) % 16)
      return false;
    if (batch > 1 && (g.strideA % EPC || g.strideB % EPC)) return false;
    const long tiles = (long)(g.M / 128) * (g.N / 128);
    if (std::is_same<TC, float>::value && tiles * batch > 512) return false;
    if (tiles >= (1L << 31) || batch > 65535) return false;
    hipLaunchKernelGGL((gemm_dma_kernel<TC, TO, BK, WNC>),
                       dim3(static_cast<unsigned>(tiles), batch), dim3(WNC * 128), 0, stream, g);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error("gemm_tn (dma): launch failed: %s", hipGetErrorString(e));
      rc = TNH_ERR_HIP;
    } else {
      rc = TNH_OK;
    }
    return true;
  } else {
    (void)g, (void)batch, (void)stream, (void)rc;
    return false;
  }
}

}  // namespace

// Tile shapes (tile_hint): 1 = 128x128 (4 waves of 64x64, K step 16 doubles / 32 floats),
// 2 = 64x64, 3 = 32x32, 4 = 16x128 (skinny M such as the thick-restart rotation),
// 5 = 128x128 with the short K step (8 doubles / 16 floats), 6 = 256x128 (8 waves),
// 7 = the direct-to-LDS 128x128 kernel with 8 waves of 64x32 (when the operands fit it).
// 0 picks by dtype and tile count (below), and the skinny tile when M <= 16.
template <typename TC, typename TA, typename TB, typename TO>
int launch_gemm_tn(const GemmArgs& g, int batch, hipStream_t stream, int tile_hint) {
  TNH_REQUIRE(g.M >= 0 && g.N >= 0 && g.K >= 0, TNH_ERR_ARG, "gemm_tn: negative size");
  int cfg = tile_hint;
  if (cfg == 0) {
    int rc = TNH_OK;
    // float64: 8 waves of 64 x 32 (two per SIMD; 0.64 vs 0.70 ms on the chi = 1024 GEMM3);
    // float32: 4 waves of 64 x 64 (8 waves measured no faster)
    constexpr int kWnc = sizeof(TC) == 8 ? 4 : 2;
    if (g.M > 0 && g.N > 0 && try_dma<TC, TA, TB, TO, kWnc>(g, batch, stream, rc)) return rc;
    const long b = batch > 0 ? batch : 1;
    const long t128 = ceil_div(g.M, 128) * ceil_div(g.N, 128) * b;
    const long t64 = ceil_div(g.M, 64) * ceil_div(g.N, 64) * b;
    // measured on MI355X at the chi = 512 / 1024 two-site shapes: 64x64 tiles lead for float64
    // (two waves per SIMD hide the load latency better than one 128x128 wave), and 128x128 with
    // the short K step leads for float32 once there are >= 4 tiles per CU
    if (g.M <= 16)
      cfg = 4;
    else if (sizeof(TC) == 4 && t128 >= 1024)
      cfg = 5;
    else if (t64 >= 256)
      cfg = 2;
    else
      cfg = 3;
  }
  constexpr bool dbl = sizeof(TC) == 8;
  constexpr int BKb = dbl ? 16 : 32;
  constexpr int BKs = dbl ? 8 : 16;
  switch (cfg) {
    case 1: return launch_one<TC, TA, TB, TO, 128, 128, BKb, 2, 2>(g, batch, stream);
    case 2: return launch_one<TC, TA, TB, TO, 64, 64, 16, 2, 2>(g, batch, stream);
    case 3: return launch_one<TC, TA, TB, TO, 32, 32, 16, 2, 2>(g, batch, stream);
    case 4: return launch_one<TC, TA, TB, TO, 16, 128, 16, 1, 4>(g, batch, stream);
    case 5: return launch_one<TC, TA, TB, TO, 128, 128, BKs, 2, 2>(g, batch, stream);
    case 6: return launch_one<TC, TA, TB, TO, 256, 128, BKs, 4, 2>(g, batch, stream);
    case 7: {
      int rc = TNH_OK;
      if (g.M > 0 && g.N > 0 && try_dma<TC, TA, TB, TO, 4>(g, batch, stream, rc)) return rc;
      return launch_one<TC, TA, TB, TO, 128, 128, BKb, 2, 2>(g, batch, stream);
    }
    default: set_error("gemm_tn: unknown tile config %d", cfg); return TNH_ERR_ARG;
  }
}

namespace {
template <typename TR, typename TA, typename TB, typename TO, int BM, int BN, int BK, int WM, int WN>
int launch_c(const GemmArgs& g, int batch, int conj_b, hipStream_t stream) {
  const long tiles = ceil_div(g.M, BM) * ceil_div(g.N, BN);
  if (tiles == 0 || batch == 0) return TNH_OK;
  TNH_REQUIRE(tiles < (1L << 31) && batch <= 65535, TNH_ERR_ARG, "gemm_tn: grid too large");
  dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(batch));
  hipLaunchKernelGGL((gemm_tn_cplx_kernel<TR, TA, TB, TO, BM, BN, BK, WM, WN>), grid,
                     dim3(WM * WN * 64), 0, stream, g, conj_b);
  TNH_CHECK_LAUNCH("gemm_tn_cplx");
  return TNH_OK;
}
}  // namespace

namespace {
// The direct-to-LDS complex kernel when the operands are stored in the compute type and the
// shape fits it (128 x 128 tiles, K a multiple of its step, 16-byte chunks: any leading
// dimension for complex128, even ones for complex64).
template <typename TR, typename TA, typename TB, typename TO>
bool try_dma_cplx(const GemmArgs& g, int batch, int conj_b, hipStream_t stream, int& rc) {
  if constexpr (std::is_same<TA, cpx<TR>>::value && std::is_same<TB, cpx<TR>>::value &&
                (std::is_same<TO, c128>::value || std::is_same<TO, c64>::value)) {
    constexpr int EPC = 16 / sizeof(TA);
    constexpr int BK = sizeof(TR) == 8 ? 8 : 16;
    if (batch < 1 || g.M <= 0 || g.N <= 0 || g.M % 128 || g.N % 128 || g.K < BK || g.K % BK)
      return false;
    if (reinterpret_cast<uintptr_t>(g.A) % 16 || reinterpret_cast<uintptr_t>(g.B) % 16 ||
        reinterpret_cast<uintptr_t>(g.C) % sizeof(TO))
      return false;
    if (g.lda % EPC || g.ldb % EPC) return false;
    if (batch > 1 && (g.strideA % EPC || g.strideB % EPC)) return false;
    const long tiles = (long)(g.M / 128) * (g.N / 128);
    if (tiles >= (1L << 31) || batch > 65535) return false;
    using TOR = typename std::conditional<std::is_same<TO, c128>::value, double, float>::type;
    hipLaunchKernelGGL((gemm_dma_cplx_kernel<TR, TOR, BK>),
                       dim3(static_cast<unsigned>(tiles), batch), dim3(512), 0, stream, g, conj_b);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
      set_error("gemm_tn (complex dma): launch failed: %s", hipGetErrorString(e));
      rc = TNH_ERR_HIP;
    } else {
      rc = TNH_OK;
    }
    return true;
  } else {
    (void)g, (void)batch, (void)conj_b, (void)stream, (void)rc;
    return false;
  }
}
}  // namespace

// Complex tiles: 1 = 64x64 (4 waves of 32x32), 2 = 32x32, 3 = 16x64 (skinny M), 4 = the
// direct-to-LDS 128x128 kernel (when the shape and storage fit it). 0 = automatic: the
// direct-to-LDS kernel when it fits, else by tile count.
template <typename TR, typename TA, typename TB, typename TO>
int launch_gemm_tn_cplx(const GemmArgs& g, int batch, int conj_b, hipStream_t stream,
                        int tile_hint) {
  TNH_REQUIRE(g.M >= 0 && g.N >= 0 && g.K >= 0, TNH_ERR_ARG, "gemm_tn: negative size");
  int cfg = tile_hint;
  if (cfg == 0 || cfg == 4) {
    int rc = TNH_OK;
    if (try_dma_cplx<TR, TA, TB, TO>(g, batch, conj_b, stream, rc)) return rc;
    if (cfg == 4) cfg = 1;
  }
  if (cfg == 0) {
    const long b = batch > 0 ? batch : 1;
    const long t64 = ceil_div(g.M, 64) * ceil_div(g.N, 64) * b;
    cfg = g.M <= 16 ? 3 : (t64 >= 256 ? 1 : 2);
  }
  constexpr int BK = sizeof(TR) == 8 ? 8 : 16;
  switch (cfg) {
    case 1: return launch_c<TR, TA, TB, TO, 64, 64, BK, 2, 2>(g, batch, conj_b, stream);
    case 2: return launch_c<TR, TA, TB, TO, 32, 32, BK, 2, 2>(g, batch, conj_b, stream);
    case 3: return launch_c<TR, TA, TB, TO, 16, 64, 16, 1, 4>(g, batch, conj_b, stream);
    default: set_error("gemm_tn: unknown complex tile config %d", cfg); return TNH_ERR_ARG;
  }
}

namespace {
template <typename TA, typename TB, typename TO, int BM, int BN, int WM, int WN, int VA, int VB>
int launch_bf16_v(const GemmArgs& g, int batch, hipStream_t stream) {
  const long tiles = ceil_div(g.M, BM) * ceil_div(g.N, BN);
  dim3 grid(static_cast<unsigned>(tiles), static_cast<unsigned>(batch));
  hipLaunchKernelGGL((gemm_tn_bf16_kernel<TA, TB, TO, BM, BN, WM, WN, VA, VB>), grid,
                     dim3(WM * WN * 64), 0, stream, g);
  TNH_CHECK_LAUNCH("gemm_tn_bf16");
  return TNH_OK;
}

template <typename TA, typename TB, typename TO, int BM, int BN, int WM, int WN>
int launch_bf16_one(const GemmArgs& g, int batch, hipStream_t stream) {
  const long tiles = ceil_div(g.M, BM) * ceil_div(g.N, BN);
  if (tiles == 0 || batch == 0) return TNH_OK;
  TNH_REQUIRE(tiles < (1L << 31) && batch <= 65535, TNH_ERR_ARG, "gemm_tn_bf16: grid too large");
  constexpr int NT = WM * WN * 64;
  constexpr int VA = 16 / sizeof(TA), VB = 16 / sizeof(TB);
  constexpr bool can_a = (32 * BM) % (VA * NT) == 0;
  constexpr bool can_b = (32 * BN) % (VB * NT) == 0;
  const bool va = can_a && vec_ok<VA, TA>(g.A, g.lda, g.M) && (batch == 1 || g.strideA % VA == 0);
  const bool vb = can_b && vec_ok<VB, TB>(g.B, g.ldb, g.N) && (batch == 1 || g.strideB % VB == 0);
  if constexpr (can_a && can_b) {
    if (va && vb) return launch_bf16_v<TA, TB, TO, BM, BN, WM, WN, VA, VB>(g, batch, stream);
  }
  if constexpr (can_a) {
    if (va) return launch_bf16_v<TA, TB, TO, BM, BN, WM, WN, VA, 1>(g, batch, stream);
  }
  if constexpr (can_b) {
    if (vb) return launch_bf16_v<TA, TB, TO, BM, BN, WM, WN, 1, VB>(g, batch, stream);
  }
  return launch_bf16_v<TA, TB, TO, BM, BN, WM, WN, 1, 1>(g, batch, stream);
}
}  // namespace

template <typename TA, typename TB, typename TO>
int launch_gemm_bf16(const GemmArgs& g, int batch, hipStream_t stream, int tile_hint) {
  TNH_REQUIRE(g.M >= 0 && g.N >= 0 && g.K >= 0, TNH_ERR_ARG, "gemm_tn_bf16: negative size");
  int cfg = tile_hint;
  if (cfg == 0) {
    const long b = batch > 0 ? batch : 1;
    cfg = ceil_div(g.M, 128) * ceil_div(g.N, 128) * b >= 512 ? 1 : 2;
  }
  switch (cfg) {
    case 1: return launch_bf16_one<TA, TB, TO, 128, 128, 2, 2>(g, batch, stream);
    case 2: return launch_bf16_one<TA, TB, TO, 64, 64, 2, 2>(g, batch, stream);
    default: set_error("gemm_tn_bf16: unknown tile config %d", cfg); return TNH_ERR_ARG;
  }
}

template int launch_gemm_bf16<__bf16, __bf16, float>(const GemmArgs&, int, hipStream_t, int);
template int launch_gemm_bf16<double, __bf16, float>(const GemmArgs&, int, hipStream_t, int);
template int launch_gemm_bf16<__bf16, __bf16, double>(const GemmArgs&, int, hipStream_t, int);

template int launch_gemm_tn_cplx<double, c128, c128, c128>(const GemmArgs&, int, int, hipStream_t, int);
template int launch_gemm_tn_cplx<float, c64, c64, c64>(const GemmArgs&, int, int, hipStream_t, int);
template int launch_gemm_tn_cplx<float, c128, c64, c64>(const GemmArgs&, int, int, hipStream_t, int);
template int launch_gemm_tn_cplx<float, c64, c64, c128>(const GemmArgs&, int, int, hipStream_t, int);

template int launch_gemm_tn<double, double, double, double>(const GemmArgs&, int, hipStream_t, int);
template int launch_gemm_tn<float, float, float, float>(const GemmArgs&, int, hipStream_t, int);
template int launch_gemm_tn<float, double, float, float>(const GemmArgs&, int, hipStream_t, int);
template int launch_gemm_tn<float, float, float, double>(const GemmArgs&, int, hipStream_t, int);

}  // namespace tnh

using namespace tnh;

extern "C" int tnh_gemm_tn(int dtype, int M, int N, int K, const void* A, long lda, const void* B,
                           long ldb, void* C, long ldc, int batch, long strideA, long strideB,
                           long strideC, int accumulate, int tile_hint, tnh_stream_t stream) {
  GemmArgs g{M, N, K, A, lda, strideA, B, ldb, strideB, C, ldc, strideC, accumulate};
  TNH_REQUIRE(lda >= M && ldb >= N && ldc >= N, TNH_ERR_ARG,
              "tnh_gemm_tn: leading dimensions (%ld, %ld, %ld) smaller than (M=%d, N=%d, N=%d)",
              lda, ldb, ldc, M, N, N);
  hipStream_t s = static_cast<hipStream_t>(stream);
  switch (dtype) {
    case TNH_F64: return launch_gemm_tn<double, double, double, double>(g, batch, s, tile_hint);
    case TNH_F32: return launch_gemm_tn<float, float, float, float>(g, batch, s, tile_hint);
    case TNH_C128:
      return launch_gemm_tn_cplx<double, c128, c128, c128>(g, batch, 0, s, tile_hint);
    case TNH_C64: return launch_gemm_tn_cplx<float, c64, c64, c64>(g, batch, 0, s, tile_hint);
    case TNH_BF16: return launch_gemm_bf16<__bf16, __bf16, float>(g, batch, s, tile_hint);
    default: set_error("tnh_gemm_tn: unsupported dtype %d", dtype); return TNH_ERR_UNSUPPORTED;
  }
}
