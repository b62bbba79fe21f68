/* libtnkit_hip -- C ABI of the MI355X (gfx950) two-site effective-Hamiltonian path.
 *
 * The reference (tnkit, Python) has no C ABI: its plug point is the duck-typed backend object
 * returned by get_backend (tnkit/backend.py:192-240) whose einsum / vdot / norm calls the
 * contraction helpers make. This library implements those contractions and the Krylov-vector
 * work behind them. Each entry cites the tnkit function it replaces.
 *
 * Conventions
 *   - Row-major C-contiguous arrays with tnkit's index order: MPS M (a, s, b); MPO W (wl, wr,
 *     s_out, s_in); environments L, R (bra, mpo, ket)  (tnkit/_dmrg.py:7-11).
 *   - Every pointer is a device pointer owned by the caller, except where marked "host".
 *     Workspaces are caller-allocated; the library never allocates device memory.
 *   - Calls are asynchronous on `stream` (a hipStream_t; NULL = the null stream) and never
 *     synchronise the host.
 *   - Every entry returns a TNH_* status; tnh_last_error() gives the message of the last
 *     failure on the calling thread.
 */
#ifndef TNKIT_HIP_H
#define TNKIT_HIP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef void* tnh_stream_t;

/* dtype codes */
enum { TNH_F64 = 0, TNH_F32 = 1, TNH_BF16 = 2, TNH_C128 = 3, TNH_C64 = 4 };
/* status codes */
enum { TNH_OK = 0, TNH_ERR_ARG = 1, TNH_ERR_WS = 2, TNH_ERR_HIP = 3, TNH_ERR_UNSUPPORTED = 4 };

/* Library version (major*10000 + minor*100 + patch) and the last error on this thread. */
int tnh_version(void);
const char* tnh_last_error(void);

/* C[m][n] (+)= sum_k A[k][m] * B[k][n], batched over `batch` with element strides.
 * The GEMM every contraction below is built from; replaces the vendor GEMMs that tnkit's
 * einsum reaches (tnkit/backend.py:317-322 numpy, torch.einsum on the torch backends).
 * tile_hint: 0 = automatic, 1 = 128x128, 2 = 64x64, 3 = 32x32, 4 = 16x128. TNH_BF16: A and B
 * in bf16 on v_mfma_f32_16x16x32_bf16, C in float32 (tiles 1 = 128x128, 2 = 64x64). */
int tnh_gemm_tn(int dtype, int M, int N, int K, const void* A, long lda, const void* B, long ldb,
                void* C, long ldc, int batch, long strideA, long strideB, long strideC,
                int accumulate, int tile_hint, tnh_stream_t stream);

/* fused_two_site (tnkit/_dmrg.py:172-180): the operator of one two-site problem, built once per
 * local solve, in the layouts the matvec streams:
 *   LT[q][p][wl] = L[p][wl][q]             (a, a, wl)
 *   S[(s1 s2 wr)][(wl s1' s2')] = sum_m W1[wl][m][s1][s1'] W2[m][wr][s2][s2']
 *   RT[wr][b][b'] = R[b'][wr][b]           (wr, b, b)
 * dtype TNH_F64 or TNH_F32; TNH_BF16 builds the bf16 operator of the BF16 matvec: L, R, LT and
 * RT in bf16, W1, W2 and S in float32. */
int tnh_fused_two_site(int dtype, int a, int wl, int wm, int wr, int d1, int d2, int b,
                       const void* L, const void* W1, const void* W2, const void* R, void* LT,
                       void* S, void* RT, tnh_stream_t stream);

/* Workspace bytes of tnh_apply_two_site_fused. */
size_t tnh_two_site_ws_bytes(int dtype, int a, int d1, int d2, int b, int wl, int wr);

/* apply_two_site_fused (tnkit/_dmrg.py:183-189): out = H_eff theta for theta (a, d1, d2, b).
 * `dtype` is the arithmetic type of the operator (F64, or F32 on the F32 matrix cores);
 * `io_dtype` the type of theta and out (F64 with an F32 operator gives the mixed-precision
 * matvec with float64 Krylov vectors; TNH_BF16 with TNH_F64 runs both GEMMs on the bf16 matrix
 * cores with float accumulation and the MPO mix in float). tile1/tile3: tile hints of the two
 * GEMMs (0 = auto). */
int tnh_apply_two_site_fused(int dtype, int io_dtype, int a, int d1, int d2, int b, int wl, int wr,
                             const void* LT, const void* S, const void* RT, const void* theta,
                             void* out, void* ws, size_t ws_bytes, int tile1, int tile3,
                             tnh_stream_t stream);

/* Bytes of the fixed device scratch of tnh_apply_two_site_fused_noalloc (192 MiB per GPU):
 * enough for a float64 operator up to a = b = 512 and a float32 operator with float64 vectors
 * up to a = b = 1024, at d = 2 and MPO bonds 5. */
#define TNH_NOALLOC_SCRATCH_BYTES (192u * 1024u * 1024u)

/* tnh_apply_two_site_fused with no workspace argument and no allocation of any kind. No
 * counterpart in tnkit: it is for callers that cannot allocate device memory (code under HIP
 * graph capture, tight serving loops, harnesses that forbid hipMalloc). The workspace is a
 * file-scope device array of TNH_NOALLOC_SCRATCH_BYTES that each GPU holds from the first call
 * on it until the process ends; shapes needing more (tnh_two_site_ws_bytes) are refused with
 * TNH_ERR_WS. All calls on one GPU share that array, so calls that overlap on one GPU (other
 * streams, other host threads) race: serialise them. */
int tnh_apply_two_site_fused_noalloc(int dtype, int io_dtype, int a, int d1, int d2, int b,
                                     int wl, int wr, const void* LT, const void* S,
                                     const void* RT, const void* theta, void* out, int tile1,
                                     int tile3, tnh_stream_t stream);

/* tnh_apply_two_site_fused restricted to some of its kernels, for timing one at a time:
 * stages bit 0 = GEMM1 (theta x L), bit 1 = the MPO mix, bit 2 = GEMM3 (x R); 7 = all. No
 * counterpart in tnkit; a profiling hook over the same workspace layout. */
int tnh_apply_two_site_stages(int dtype, int io_dtype, int a, int d1, int d2, int b, int wl,
                              int wr, const void* LT, const void* S, const void* RT,
                              const void* theta, void* out, void* ws, size_t ws_bytes, int tile1,
                              int tile3, int stages, tnh_stream_t stream);

/* Workspace bytes of the environment updates, for bonds (a, c), physical d, MPO bonds wl, wr. */
size_t tnh_env_ws_bytes(int dtype, int a, int d, int c, int wl, int wr);

/* update_left_env (tnkit/_dmrg.py:216-222): out (c, wr, c) = L (a, wl, a) grown past A (a, d, c)
 * with W (wl, wr, d, d). dtypes F64, F32, C128, C64 (complex: the bra is conj(A)). */
int tnh_update_left_env(int dtype, int a, int d, int c, int wl, int wr, const void* L,
                        const void* A, const void* W, void* out, void* ws, size_t ws_bytes,
                        tnh_stream_t stream);

/* update_right_env (tnkit/_dmrg.py:225-231): out (c, wl, c) = R (b, wr, b) grown past
 * B (c, d, b) with W (wl, wr, d, d). dtypes F64, F32, C128, C64 (complex: the bra is conj(B)). */
int tnh_update_right_env(int dtype, int c, int d, int b, int wl, int wr, const void* R,
                         const void* B, const void* W, void* out, void* ws, size_t ws_bytes,
                         tnh_stream_t stream);

/* ---- backward (no tnkit counterpart: tnkit has no autograd; these are the hand-written
 * backward kernels of the differentiation study) -------------------------------------------- */

/* Workspace bytes of tnh_matvec2_bwd. */
size_t tnh_matvec2_bwd_ws_bytes(int dtype, int a, int d1, int d2, int b, int wl, int wr);

/* The VJP of apply_two_site_fused (tnkit/_dmrg.py:183-189) with respect to theta:
 * theta_bar (a, d1, d2, b) = H_eff^H ybar, H_eff built from L (a, wl, a), W1 (wl, wm, d1, d1),
 * W2 (wm, wr, d2, d2), R (b, wr, b). The adjoint operator (bra and ket of L and R swapped, s and
 * s' of W swapped, all conjugated) is built in the workspace and applied by the forward kernels.
 * dtypes F64, F32, C128, C64. */
int tnh_matvec2_bwd(int dtype, int a, int d1, int d2, int b, int wl, int wm, int wr,
                    const void* L, const void* W1, const void* W2, const void* R,
                    const void* ybar, void* theta_bar, void* ws, size_t ws_bytes,
                    tnh_stream_t stream);

/* Workspace bytes of tnh_env_left_bwd / tnh_env_right_bwd for bonds (a, c). */
size_t tnh_env_bwd_ws_bytes(int dtype, int a, int d, int c, int wl, int wr);

/* The VJP of update_left_env (tnkit/_dmrg.py:216-222) with respect to L: given the cotangent
 * Lbar_out (c, wr, c) of L' = L grown past A (a, d, c) with W (wl, wr, d, d), writes
 * Lbar (a, wl, a) (torch's convention: the conjugate of dF/dL for F = sum conj(Lbar_out) L').
 * It is the right-environment update of conj(Lbar_out) past A, conjugated. */
int tnh_env_left_bwd(int dtype, int a, int d, int c, int wl, int wr, const void* Lbar_out,
                     const void* A, const void* W, void* Lbar, void* ws, size_t ws_bytes,
                     tnh_stream_t stream);

/* The VJP of update_right_env (tnkit/_dmrg.py:225-231) with respect to R: Rbar_out (c, wl, c),
 * B (c, d, b), W (wl, wr, d, d) -> Rbar (b, wr, b); the mirror of tnh_env_left_bwd. */
int tnh_env_right_bwd(int dtype, int c, int d, int b, int wl, int wr, const void* Rbar_out,
                      const void* B, const void* W, void* Rbar, void* ws, size_t ws_bytes,
                      tnh_stream_t stream);

/* ---- hipGraph capture of library calls (replaces tnkit/block_graph.py:89-180 and the capture
 * half of tnkit/cudagraph.py:87-342 for C callers): tnh_graph_begin starts a thread-local
 * capture on `stream` (not the null stream); every library call issued on it is recorded (the
 * library allocates nothing and never synchronises) except the cooperative solvers tnh_sytrd,
 * tnh_syev_top, tnh_heev_top and tnh_qr, which are for eager streams only: each resets its grid
 * barrier words before every cooperative launch (tnh_heev_top and tnh_qr with hipMemsetAsync,
 * whose captured nodes are not replay-safe), so do not issue them inside a capture; tnh_graph_end
 * instantiates the recording into *graph; tnh_graph_launch replays it on a stream;
 * tnh_graph_destroy frees it. Pointers are baked in: replays read and write the same buffers. */
int tnh_graph_begin(tnh_stream_t stream);
int tnh_graph_end(tnh_stream_t stream, void** graph);
int tnh_graph_launch(void* graph, tnh_stream_t stream);
int tnh_graph_destroy(void* graph);

/* ---- batched forms: `batch` independent problems of one shape, element strides s* between
 * consecutive problems (the legs of tnkit's batched evolution, tnkit/batch_tdvp.py:211-243).
 * Element types F64, F32, C128, C64; workspaces are `batch` times the single-problem size. */
int tnh_fused_two_site_batched(int dtype, int batch, int a, int wl, int wm, int wr, int d1, int d2,
                               int b, const void* L, long sL, const void* W1, long sW1,
                               const void* W2, long sW2, const void* R, long sR, void* LT,
                               long sLT, void* S, long sS, void* RT, long sRT, tnh_stream_t stream);

/* apply_two_site_fused over a batch (tnkit/batch_tdvp.py:211-219 apply_two_site_b). io_dtype may
 * be C128 with a C64 operator (complex64 matvec with complex128 Krylov vectors). */
int tnh_apply_two_site_batched(int dtype, int io_dtype, int batch, int a, int d1, int d2, int b,
                               int wl, int wr, const void* LT, long sLT, const void* S, long sS,
                               const void* RT, long sRT, const void* theta, long sth, void* out,
                               long sout, void* ws, size_t ws_bytes, int tile1, int tile3,
                               int stages, tnh_stream_t stream);

/* The batched matvec over legs of different bonds held in one padded batch (tnkit's ragged
 * batches, tnkit/batch_tdvp.py:42-55 and :572-681, where every leg computes at the stacked
 * width): live[2z], live[2z+1] (device ints) = leg z's live left and right bonds (a_z <= a,
 * b_z <= b; theta, L and R are zero outside them, as the masked batch holds them). Same operands
 * and workspace as tnh_apply_two_site_batched with all three stages; complex128 vectors with a
 * complex128 or complex64 operator. On shapes that take the direct-to-LDS GEMMs (stacked
 * d1 d2 b, a wl, a d1 d2 and b multiples of 128) the tiles of a leg that lie outside its live
 * block skip their products and store zeros, and the K loops stop at its live extent, so a leg
 * costs about its own bonds rather than the stacked ones; other shapes compute in full. The
 * output is zero outside every leg's live block either way. */
int tnh_apply_two_site_ragged(int dtype, int io_dtype, int batch, int a, int d1, int d2, int b,
                              int wl, int wr, const void* LT, long sLT, const void* S, long sS,
                              const void* RT, long sRT, const void* theta, long sth, void* out,
                              long sout, const int* live, void* ws, size_t ws_bytes,
                              tnh_stream_t stream);

/* update_left_env / update_right_env over a batch (tnkit/batch_tdvp.py:231-243); complex inputs
 * take conj(A) / conj(B) as tnkit does. */
int tnh_update_left_env_batched(int dtype, int batch, int a, int d, int c, int wl, int wr,
                                const void* L, long sL, const void* A, long sA, const void* W,
                                long sW, void* out, long sout, void* ws, size_t ws_bytes,
                                tnh_stream_t stream);
int tnh_update_right_env_batched(int dtype, int batch, int c, int d, int b, int wl, int wr,
                                 const void* R, long sR, const void* B, long sB, const void* W,
                                 long sW, void* out, long sout, void* ws, size_t ws_bytes,
                                 tnh_stream_t stream);

/* ---- complex128 Krylov exponential, batched over legs (tnkit/eigensolver.py:95-217,
 * tnkit/batch_tdvp.py:247-324). Per leg: basis V (m rows of ldv, leg stride sV), work vector w
 * (leg stride sw), and ab[leg*sab + k] = alpha_k, ab[leg*sab + m + k] = beta_k (device). */
size_t tnh_expm_ws_bytes(int batch);

/* Step j of the fixed-m Lanczos exponential after w = H V[j]: three-term subtraction with
 * beta_{j-1}, CGS2 against V[0..j] with alpha_j read from the first projection, beta_j = |w|,
 * V[j+1] = w / max(beta_j, 1e-200). Nothing is read back to the host. m <= 32. */
int tnh_expm_step_c(int batch, long n, int m, int j, void* V, long ldv, long sV, void* w, long sw,
                    double* ab, long sab, void* ws, size_t ws_bytes, tnh_stream_t stream);

/* tnh_expm_step_c over a padded batch of legs with different bonds (tnkit's ragged batches,
 * tnkit/batch_tdvp.py:42-55): live[2z], live[2z+1] (device ints) = leg z's live extents (x, y) in
 * vectors laid out as rows of `rowlen` entries whose last index runs over `bpad` (theta (a, d1,
 * d2, b): rowlen = d1 d2 b, bpad = b; x, y = the leg's left and right bonds). Entry i of a leg is
 * live when i / rowlen < x and i % bpad < y; the others are zero in every vector of the leg and
 * the passes skip them, with the same result. */
int tnh_expm_step_c_ragged(int batch, long n, int m, int j, void* V, long ldv, long sV, void* w,
                           long sw, double* ab, long sab, const int* live, int rowlen, int bpad,
                           void* ws, size_t ws_bytes, tnh_stream_t stream);

/* out[leg] = sum_{k<nrows} c[leg][k] V[leg][k] with complex coefficients c (device, interleaved
 * re/im, leg stride sc doubles): the final combination beta0 * expm(tau T)[:, 0] . V. */
int tnh_lincomb_c(int batch, long n, int nrows, const void* V, long ldv, long sV, const double* c,
                  long sc, void* out, long sout, tnh_stream_t stream);
/* tnh_lincomb_c over a padded batch of legs with different bonds: live, rowlen and bpad as for
 * tnh_expm_step_c_ragged; a leg's entries outside its live extents are written as zeros without
 * reading V (they are zero in V). */
int tnh_lincomb_c_ragged(int batch, long n, int nrows, const void* V, long ldv, long sV,
                         const double* c, long sc, void* out, long sout, const int* live,
                         int rowlen, int bpad, tnh_stream_t stream);

/* Workspace bytes of the Krylov-vector entries below. */
size_t tnh_krylov_ws_bytes(long n, int nmax);

/* One thick-restart Lanczos orthogonalisation step, float64 (tnkit/eigensolver.py:560-604):
 *   w -= sum_{r<nsub} sub_coef[r] V[sub_base + r]      (three-term or arrowhead coupling;
 *                                                       sub_coef is a HOST array)
 *   scal[0] = alpha = <V[j], w>;  w -= alpha V[j]
 *   two classical Gram-Schmidt passes against V[0..nlive)
 *   scal[1] = beta = |w|
 * V is (rows, ldv) row-major; nlive <= 64. scal is device memory (2 doubles). */
int tnh_krylov_orth(long n, int nlive, const double* V, long ldv, double* w, int j, int nsub,
                    int sub_base, const double* sub_coef, double* scal, void* ws, size_t ws_bytes,
                    tnh_stream_t stream);

/* y += weight * sum_{r<k} <V[r], x> V[r], float64, k <= 64, coefficients reduced on the device
 * (no host read). x may alias y. With weight = penalty and V the block vectors of earlier states
 * it is the orthogonality penalty of tnkit/dmrg_excited.py:709-720; with weight = -1 and V
 * orthonormal it is the projector off span(V) of tnkit/dmrg_excited.py:384-397. The workspace
 * is tnh_krylov_ws_bytes. */
int tnh_rank_update(long n, int k, const double* V, long ldv, const double* x, double weight,
                    double* y, void* ws, size_t ws_bytes, tnh_stream_t stream);

/* out = beta * out + sum_{r<nrows} coef[r] V[r]  (coef HOST, nrows < 128): the Ritz vector
 * reconstruction of tnkit/eigensolver.py:670-673 and the Krylov combination of
 * tnkit/eigensolver.py:212. */
int tnh_lincomb(long n, int nrows, const double* V, long ldv, const double* coef, double beta,
                double* out, tnh_stream_t stream);

/* out[0] = |x|  (tnkit NumpyBackend.norm, backend.py:295-296), float64, deterministic. */
int tnh_norm(long n, const double* x, double* out, void* ws, size_t ws_bytes, tnh_stream_t stream);

/* out[0] = <x, y>  (tnkit NumpyBackend.vdot for real data, backend.py:292-293). */
int tnh_dot(long n, const double* x, const double* y, double* out, void* ws, size_t ws_bytes,
            tnh_stream_t stream);

/* Workspace bytes of tnh_rows_dot. */
size_t tnh_rows_ws_bytes(int dtype);

/* out[r] = sum_i V[r*ldv + i] x[i] for r < k (no conjugation): the einsum "kn,n->k" of tnkit's
 * Krylov projections (tnkit/eigensolver.py:160, :588, :594) as the backend receives it. dtypes
 * F64, F32, C128, C64; accumulates in float64 / complex128; deterministic. */
int tnh_rows_dot(int dtype, int k, long n, const void* V, long ldv, const void* x, void* out,
                 void* ws, size_t ws_bytes, tnh_stream_t stream);

/* out[i] = sum_{r<k} c[r] V[r*ldv + i] with c on the device: the einsums "k,kn->n",
 * "r,rn->n", "j,jn->n" of tnkit's eigensolvers (tnkit/eigensolver.py:161, :567, :589, :595,
 * :672). out may not alias V. */
int tnh_rows_combine(int dtype, int k, long n, const void* c, const void* V, long ldv, void* out,
                     tnh_stream_t stream);

/* y = x * (mul / max(div[0], 1e-200)); div may be NULL (then y = mul * x). The division guard is
 * tnkit's guard_denom (tnkit/backend.py:699-708). */
int tnh_scale_by(long n, const double* x, const double* div, double mul, double* y,
                 tnh_stream_t stream);

/* dst = (dst_dtype) src, for F64 <-> F32 and F64 -> BF16 (round to nearest even). */
int tnh_convert(int dst_dtype, int src_dtype, long n, const void* src, void* dst,
                tnh_stream_t stream);

/* ---- U(1) block-sparse contractions -------------------------------------------------------
 * A block tensor is one flat buffer of blocks; the host builds its layouts and contraction plans
 * (tnkit/blocks.py:220-253, :952-1100). */

/* Output elements handled per workgroup by tnh_gather_runs (the chunk of its chunk table). */
int tnh_gather_chunk(void);

/* Replaces tnkit's gather _Index.take (tnkit/blocks.py:363-495, used by tensordot :1168-1169 and
 * BlockTensor.canonical :755-760): dst[j] = (dst_dtype) src[sbeg[r] + (j - dbeg[r]) * sstep[r]]
 * for the run r covering j; where no run covers j, dst[j] = 0 if `fill`, else untouched. `runs`
 * (device) holds four arrays of nruns longs: sbeg, sstep, dbeg, len, sorted by dbeg with
 * disjoint runs. `chunks` (device) holds 2 * nchunks ints, nchunks = ceil(n /
 * tnh_gather_chunk()): the first run that may overlap chunk c, then the end of the runs that do.
 * Dtypes F64/F32 in any combination. */
int tnh_gather_runs(int src_dtype, int dst_dtype, long n, const void* src, void* dst, long nruns,
                    const long* runs, const int* chunks, int fill, tnh_stream_t stream);

/* The 2-D part of the same gathers (transposed blocks, and blocks of many short runs): tile t
 * of `tiles` (device, ntiles x 7 longs: src0, srow, scol, dst0, drow, nrows <= 64,
 * ncols <= 64) moves dst[dst0 + i*drow + j] = (dst_dtype) src[src0 + i*srow + j*scol];
 * transposes (scol != 1) go through LDS. */
int tnh_gather_tiles(int src_dtype, int dst_dtype, long ntiles, const void* src, void* dst,
                     const long* tiles, tnh_stream_t stream);

/* Tile shape (bm x bn) of grouped-GEMM tile config `cfg` (1 = 128x128, 2 = 64x64, 3 = 32x32,
 * 4 = 16x128, 5 = 64x128) and, when max_grid is not NULL, the persistent grid for it on the
 * current device (resident workgroups on all CUs, a multiple of 8). */
int tnh_grouped_tile(int dtype, int cfg, int* bm, int* bn, int* max_grid);

/* All charge groups of one block contraction in one persistent launch; replaces the per-charge
 * matmul loop of tnkit's tensordot (tnkit/blocks.py:1171-1190). Group g, descriptor
 * desc[8g..8g+8) = (qa, lda, qb, ldb, po, R, K, C) (device longs):
 *   C[po + r*C + c] = sum_{k<K} A[qa + k*lda + r] * B[qb + k*ldb + c],  r < R, c < C  (K = 0: 0)
 * A and B are 16-byte aligned K-major stacks whose rows are padded with zeros to lda, ldb (both
 * multiples of 4, as are qa and qb). tiles (device, ntiles x 3 ints) = (group, tile row, tile
 * column) in config `tile_cfg`'s tile shape; workgroup b of `grid` (a multiple of 8) runs
 * positions b, b + grid, ...; a group < 0 marks an empty position. dtype F64 or F32. */
int tnh_grouped_gemm(int dtype, const long* desc, long ntiles, const int* tiles, const void* A,
                     const void* B, void* C, int tile_cfg, int grid, tnh_stream_t stream);

/* The same grouped product when every group has R <= 8 and K <= 8 (the MPO contractions of the
 * two-site matvec, tnkit/_dmrg.py:166-167): memory bound, one thread per output column.
 * chunks (device, nchunks x 2 ints) = (group, first column) of each 1024-column chunk. */
int tnh_grouped_small(int dtype, const long* desc, long nchunks, const int* chunks, const void* A,
                      const void* B, void* C, tnh_stream_t stream);

/* The same grouped product when every group has R <= 32 and K <= 32: the two MPO tensors of
 * the two-site matvec (tnkit/_dmrg.py:166-167) contracted into one operator applied in one
 * pass, whose charge groups have up to D d^2 rows and columns. Same descriptors and chunk table
 * as tnh_grouped_small; the group's A is staged in LDS. */
int tnh_grouped_mid(int dtype, const long* desc, long nchunks, const int* chunks, const void* A,
                    const void* B, void* C, tnh_stream_t stream);

/* Batched eigendecomposition of small Hermitian (TNH_C128) or real symmetric (TNH_F64) matrices,
 * 1 <= n <= 64: the eigh of tnkit's Gram truncation (tnkit/linalg.py:149, batched over the legs
 * in tnkit/batch_tdvp.py:572-599). Matrix b is A + b*sA with row i at i*lda (both triangles read),
 * plus add[b*sadd + j] on the diagonal when add is not NULL (real, may be NULL). Writes the
 * eigenvalues ascending to w + b*sw and the eigenvectors as columns, Z[b*sZ + i*ldz + j] =
 * component i of eigenvector j; Z may alias A. One workgroup per matrix, cyclic Jacobi in LDS;
 * a matrix stops after a sweep without rotations or with off-diagonal norm <= tol * ||A||_F.
 * sweeps (device ints, may be NULL): sweeps run per matrix, negative when max_sweeps ended it
 * unconverged or the input was not finite. */
int tnh_eigh_jacobi(int dtype, long batch, int n, const void* A, long lda, long sA,
                    const double* add, long sadd, double* w, long sw, void* Z, long ldz, long sZ,
                    int max_sweeps, double tol, int* sweeps, tnh_stream_t stream);

/* Thin SVD A = U diag(S) Vh of small float64 (TNH_F64) or complex128 (TNH_C128) matrices with
 * min(m, n) <= 64 and tnh_svd_jacobi_lds_bytes(dtype, m, n) <= 162816: the economy SVD of the
 * backends (tnkit/backend.py:441-447, :784-800) that the 'svd' truncation calls
 * (tnkit/_dmrg.py:116-155). Matrix b is A + b*sA with row i at i*lda; with k = min(m, n) it
 * writes S + b*sS (k values, descending), U + b*sU (m x k, row stride ldu) and Vh + b*sV
 * (k x n, row stride ldv). One workgroup per matrix, one-sided Jacobi in LDS; exactly zero
 * singular values get an orthonormal completion, so U and Vh are always isometries. A matrix
 * stops after a sweep without rotations (|b_p^H b_q| <= tol |b_p| |b_q| for every column pair).
 * sweeps (device ints, may be NULL): sweeps run per matrix, negative when max_sweeps ended it
 * or the input was not finite. */
size_t tnh_svd_jacobi_lds_bytes(int dtype, int m, int n);
int tnh_svd_jacobi(int dtype, long batch, int m, int n, const void* A, long lda, long sA, double* S,
                   long sS, void* U, long ldu, long sU, void* Vh, long ldv, long sV, int max_sweeps,
                   double tol, int* sweeps, tnh_stream_t stream);

/* Householder tridiagonalisation A = Q T Q^T of a real symmetric float64 matrix, 1 <= n <= 2048
 * (the first phase of the Gram eigendecomposition, tnkit/linalg.py:149). One cooperative launch
 * of ceil(n / 8) workgroups, row i of A held by workgroup i mod G. Row i of A is read at
 * A + i*lda as given (pass a symmetric matrix). Writes d (n diagonal), e (n - 1 subdiagonal),
 * tau (n - 1) and Vr (n x n, row j = reflector v_j: zero at and above j, 1 at j + 1), with
 * Q = H_0 ... H_{n-2}, H_j = I - tau_j v_j v_j^T. status (device int, may be NULL): 0, or 1 when
 * a workgroup's bounded wait ran out (the result is then invalid). */
size_t tnh_sytrd_ws_bytes(int n);
int tnh_sytrd(int n, const double* A, long lda, double* d, double* e, double* tau, double* Vr,
              int* status, void* ws, size_t ws_bytes, tnh_stream_t stream);

/* All eigenvalues and the k leading eigenvectors of a real symmetric float64 matrix,
 * 2 <= n <= 2048: the eigh of tnkit's Gram truncation at orders above 64
 * (tnkit/linalg.py:138-159, `be.eigh(A @ A^H)` at :149; the block per-sector form
 * tnkit/blocks.py:1345-1377). tnh_sytrd, divide and conquer on the tridiagonal (leaves of order
 * <= 32 on tnh_eigh_jacobi), and the WY back-transformation of the k vectors on tnh_gemm_tn.
 * Writes w (n eigenvalues, ascending) and Z (n x k, row stride ldz): column c = eigenvector of
 * the c-th largest eigenvalue. info (device int, may be NULL): 0; bit 0: a wait of the
 * tridiagonalisation ran out; bit 1: a leaf eigensolve did not converge. */
size_t tnh_syev_top_ws_bytes(int n, int k);
int tnh_syev_top(int n, int k, const double* A, long lda, double* w, double* Z, long ldz,
                 int* info, void* ws, size_t ws_bytes, tnh_stream_t stream);

/* All eigenvalues and the k leading eigenvectors of a batch of complex Hermitian (complex128)
 * matrices, 2 <= n <= 1024: the complex eigh of tnkit's Gram truncation at orders above 64
 * (tnkit/linalg.py:138-159, `be.eigh(A @ A^H)` at :149), batched over the legs of the batched
 * TDVP (tnkit/batch_tdvp.py:572-599, one torch.linalg.eigh over the stacked Gram matrices).
 * Matrix m is A + m*sA (complex elements, row i at i*lda, both triangles read as given: pass a
 * Hermitian matrix). Householder reduction to a real tridiagonal T (LAPACK zhetd2's
 * reflectors, one cooperative launch per group of resident matrices), the divide and conquer of
 * tnh_syev_top on T, and the complex WY back-transformation. Writes w + m*sw (n eigenvalues,
 * ascending) and Z + m*sZ (n x k complex, row stride ldz): column c = eigenvector of the c-th
 * largest eigenvalue. info (device int, may be NULL): 0; bit 0: a wait of a reduction ran out;
 * bit 1: a leaf eigensolve did not converge (over the whole batch). The workspace size depends on
 * the device (how many matrices one launch holds): query it on the device of the call. */
size_t tnh_heev_top_ws_bytes(int batch, int n, int k);
int tnh_heev_top(int batch, int n, int k, const void* A, long lda, long sA, double* w, long sw,
                 void* Z, long ldz, long sZ, int* info, void* ws, size_t ws_bytes,
                 tnh_stream_t stream);

/* Thin QR A = Q R of a batch of tall float64 (TNH_F64, m <= 2048) or complex128 (TNH_C128,
 * m <= 1024) matrices, 1 <= n <= m: the QR that re-orthonormalises the projected factor A Z of
 * tnkit's Gram truncation when the Gram matrix is taken on the short side (tnkit/linalg.py:155-159,
 * `Q, R = be.qr(A @ Z)`; tnkit/blocks.py:1345-1377 per sector). Matrix q is A + q*sA (m x n,
 * row stride lda). LAPACK geqrf's reflectors (R with a real diagonal, the R geqrf returns) in one
 * cooperative launch per group of resident matrices, Q = H_0 ... H_{n-1} [I; 0] (ungqr's Q) by
 * WY blocks on tnh_gemm_tn. Writes Q + q*sQ (m x n, row stride ldq) and R + q*sR (n x n upper,
 * zeros below, row stride ldr). info (device int, may be NULL): 0, or 1 when a bounded wait ran
 * out. The workspace size depends on the device: query it on the device of the call. */
size_t tnh_qr_ws_bytes(int dtype, int batch, int m, int n);
int tnh_qr(int dtype, int batch, int m, int n, const void* A, long lda, long sA, void* Q, long ldq,
           long sQ, void* R, long ldr, long sR, int* info, void* ws, size_t ws_bytes,
           tnh_stream_t stream);

#ifdef __cplusplus
}
#endif

#endif /* TNKIT_HIP_H */
