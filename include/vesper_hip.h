/*
 * vesper_hip.h -- C-ABI of libvesper_hip.so, the MI355X (gfx950) map-alignment
 * hot path of VESPER_Python.
 *
 * Conventions
 *   - Every pointer named *_dev is device memory owned by the caller; *_host
 *     is host memory.  Plans own their workspaces; vx_score_rows and
 *     vx_topn_ws allocate nothing.  vx_topn, vx_map_stats, vx_get_score and
 *     vx_threshold_dmax share a library-owned scratch that grows on demand
 *     (hipDeviceSynchronize + hipFree + hipMalloc when a call needs more than
 *     every earlier one): never record them into a HIP graph.
 *   - All work is enqueued on the caller's stream (a hipStream_t passed as
 *     void*; NULL = the null stream).  Calls that return host values
 *     (vx_map_stats, vx_get_score, vx_plan_info) synchronise that stream.
 *   - Status: 0 = OK, negative = error; vx_last_error() returns the message of
 *     the last failure on the calling thread.  Nothing aborts.
 *   - One host thread drives a plan at a time.
 *   - Arrays are C-ordered [x][y][z] as the reference indexes them after its
 *     swapaxes (VESPER_1_prob_new.py:43); vectors are [x][y][z][3];
 *     probability channels are 4 planes [4][x][y][z].
 *   - Rotation rows are float32 3x3 matrices F = float32(flip(M)), row-major,
 *     where M = R.from_euler("ZYX", angle, degrees=True).as_matrix() with
 *     |m| <= 1e-15 zeroed (VESPER_1_prob_new.py:305-307).  Positions use F^T
 *     (:310), vectors use F (:341).
 */
#ifndef VESPER_HIP_H
#define VESPER_HIP_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define VX_ABI_VERSION 1

/* status codes */
#define VX_OK 0
#define VX_ERR_ARG (-1)
#define VX_ERR_HIP (-2)
#define VX_ERR_UNSUPPORTED (-3)
#define VX_ERR_NOMEM (-4)
#define VX_ERR_STATE (-5)

/* search modes: the -M letters V/O/C/P/L (main.py:194-203) and prob mode */
#define VX_MODE_VEC 0
#define VX_MODE_OVERLAP 1
#define VX_MODE_CC 2
#define VX_MODE_PCC 3
#define VX_MODE_LAPLACIAN 4
#define VX_MODE_PROB 5

/* flags */
#define VX_FLAG_NO_FMA 1   /* K=3 rotation products rounded separately (no FMA) */
#define VX_FLAG_NO_PRUNE 2 /* transform the whole grid instead of the query's support box */

typedef struct vx_plan vx_plan;
typedef void* vx_stream; /* hipStream_t */

/* ABI version (VX_ABI_VERSION). */
int vx_abi_version(void);

/* Message of the last failing call on this thread ("" if none). */
const char* vx_last_error(void);

/*
 * Gaussian-weighted vector field.  Replaces doFastVEC/calc
 * (VESPER_1_prob_new.py:180-257, :121-142), called by fastVEC (:145-177).
 * src_dev: thresholded source density [sx][sy][sz] float32; geometry in
 * float64 (src_width = the float32 voxel width, src_orig3 = the float32 header
 * origin, both widened).  Writes data_dev [n][n][n] float64 and vec_dev
 * [n][n][n][3] float32 (zero where the reference skips a voxel).
 */
int vx_fastvec(const float* src_dev, int sx, int sy, int sz, double src_width,
               const double* src_orig3_host, int n, double dest_width,
               const double* dest_orig3_host, double dreso, double* data_dev,
               float* vec_dev, vx_stream stream);

/*
 * Map statistics after fastVEC (VESPER_1_prob_new.py:157-175):
 * out5_host = {sum, count_nonzero, mean(>0), ||>0||_2, ||>0 - mean||_2}.
 */
int vx_map_stats(const double* data_dev, int64_t count, double* out5_host,
                 vx_stream stream);

/*
 * Search plan for grid edge n, a mode, and `batch` rotations scored per
 * launch.  Replaces the per-rotation pyFFTW plans and buffers of
 * rot_and_search_fft (VESPER_1_prob_new.py:1422-1428, :1503-1509).
 */
int vx_plan_create(int n, int mode, int batch, int flags, vx_plan** out);
int vx_plan_destroy(vx_plan* plan);

/*
 * Kernel schedule of a plan (the autotuner's knobs).  VX_TUNE_XPASS_VARIANT:
 * 0 = FFT x pass accumulating channels in registers, 1 = in LDS, 3 = in
 * registers with the next channel prefetched (tiles of N x tile <= 2048; 0,
 * 1 and 3 give identical results), 2 = split-bf16 MFMA DFT x pass (one
 * correlation group, N <= 192; scores within 1e-4 relative of the FFT pass),
 * 4 = radix-split x pass: N = N1 x N2 (N1 <= 32, N2 in 2, 3, 4, 5, 7, 8), the
 * length-N1 DFTs on split-f16 MFMA with per-tile power-of-two scaling, the
 * radix-N2 stage in fp32 registers (one correlation group; scores within
 * 1e-5 relative of the FFT pass).  vx_rescore_f64 always takes its
 * candidates from the FFT x pass.  VX_TUNE_XPASS_TILE: columns per x-pass
 * tile (0 = default; 32 or 64 for variant 2; variant 4 uses 16).  VX_TUNE_SLAB_THREADS: threads per z/y slab workgroup (0 =
 * default = 1024 where the slab fits that shape, else 512; identical results).
 * Grids whose (y, z) slab exceeds LDS always use the pencil passes.
 */
#define VX_TUNE_XPASS_VARIANT 1
#define VX_TUNE_XPASS_TILE 2
#define VX_TUNE_FORCE_PENCIL 3 /* 1: separate z and y passes even where a slab fits LDS */
#define VX_TUNE_PIN_BOX_WIDTH 4 /* > 0: fix the support box to [PIN_BOX_LO, PIN_BOX_LO + width) so
                                   launch shapes stay constant across queries (captured graphs);
                                   vx_set_search fails if a query's support does not fit it;
                                   0 = per query.  Set VX_TUNE_PIN_BOX_LO first. */
#define VX_TUNE_PIN_BOX_LO 5
#define VX_TUNE_SLAB_THREADS 6
#define VX_TUNE_STREAMS 7 /* 2: two batches in flight on two workspaces and streams (the
                             second stream is forked from and joined to the caller's);
                             1 = one (default).  Identical results. */
#define VX_TUNE_STATIC_PLANS 8 /* 1 (default): grid edges with compile-time FFT stage plans
                                  (every FFTW-fast length from 16 to 200) run kernels with
                                  every stage's radix, span and twiddle offset as constants;
                                  0 = runtime descriptors only.  Results equal to fp32
                                  rounding (bit-identical at 175). */
#define VX_TUNE_RADIX_ORDER 9 /* order of the FFT stages: 0 (default) = radix 8, 4, 2 then the odd
                                 primes ascending; 1 = reversed; 2 = odd primes ascending first;
                                 3 = odd primes descending first.  Results equal to fp32 rounding. */
#define VX_TUNE_SLAB_FUSED 10 /* 1 (default): on a compile-time plan the (y, z) slab kernels read
                                 their first stage from HBM, fold the half-spectrum unpack / pack
                                 into the stage after it and write (or reduce) their last stage
                                 from registers; 0 = separate load, unpack/pack and store passes
                                 through LDS.  Results equal to fp32 rounding (bit-identical
                                 at 175). */
#define VX_TUNE_PENCIL_CHUNKS 11 /* grid edges 201..400 on the fused pencil passes: chunks of
                                    line pairs each workgroup of the inverse z pass walks (one
                                    argmax atomic per workgroup); 0 (default) = automatic.
                                    Bit-identical results for every value. */
int vx_plan_tune(vx_plan* plan, int key, int value);

/* info_host[8] = {n, mode, batch, groups, query channels, workspace bytes,
 *                 support box start, support box width} */
int vx_plan_info(const vx_plan* plan, int64_t* info_host);

/*
 * Target spectra: conj(rfftn(target channel)) for the plan's mode
 * (VESPER_1_prob_new.py:600-628; prob mode :829-864).
 * tdata_dev float64 [n^3]; tvec_dev float32 [n^3 x 3]; tprob_dev float32
 * [4][n^3] (prob mode only, else NULL); t_ave = target.ave (PCC mode).
 */
int vx_set_target(vx_plan* plan, const double* tdata_dev, const float* tvec_dev,
                  const float* tprob_dev, double t_ave, vx_stream stream);

/*
 * Query map to rotate (the search map's resampled data/vec/prob), and the
 * target average used by PCC query channels (VESPER_1_prob_new.py:1450).
 */
int vx_set_search(vx_plan* plan, const double* sdata_dev, const float* svec_dev,
                  const float* sprob_dev, double t_ave, vx_stream stream);

/*
 * Prob mode: enable the mixed score of find_best_trans_mixed
 * (VESPER_1_prob_new.py:1530-1545) with the pass-1 statistics.
 */
int vx_set_prob_stats(vx_plan* plan, int enable, double alpha, double vstd,
                      double vave, double pstd, double pave);

/*
 * Score `nrows` rotations: rot_and_search_fft / rot_and_search_fft_prob
 * (VESPER_1_prob_new.py:1418-1527) for every row of rot_dev [nrows][9].
 * Per row and group g (G = 1; prob mode: G = 2 {vec, prob}, or 3 {vec, prob,
 * mixed} once vx_set_prob_stats enabled mixing):
 *   best_dev[row*G + g]          float32 max of the channel-summed,
 *                                unnormalised circular correlation (N^3 x the
 *                                reference's score before CC/PCC scaling)
 *   trans_dev[(row*G + g)*3 + a] its first argmax in C order (x, y, z)
 */
int vx_score_rows(vx_plan* plan, const float* rot_dev, int nrows, float* best_dev,
                  int32_t* trans_dev, vx_stream stream);

/*
 * FP64 re-scoring of rows (north-star precision path, no counterpart in the
 * reference, which ranks by its float32 maxima, VESPER_1_prob_new.py:664,
 * :739): for every row of rot_dev [nrows][9], each translation whose float32
 * correlation (the search's own, per output group; prob mode's mixed score as
 * the float32 search forms it) is >= best - window * |best| is a candidate,
 * and its correlation N^3 sum_v T[v] Q[(v + t) mod N] is summed directly in
 * FP64: target channels in float64 from tdata/tvec/tprob (the mode's
 * transform, t_ave for PCC), the rotated query with float32 positions and
 * FP64-rotated vectors from sdata/svec/sprob -- the same maps as
 * vx_set_target / vx_set_search.  Writes, per row and group (layout of
 * vx_score_rows; prob mode's mixed score with the vx_set_prob_stats
 * statistics in FP64):
 *   score_dev[row*G + g]           FP64 value of the best candidate (largest,
 *                                  then smallest C-order index)
 *   trans_dev[(row*G + g)*3 + a]   its translation
 *   ncand_dev[row*G + g]           candidates found; > max_cand means only
 *                                  max_cand of them were summed and the row
 *                                  needs a smaller window or a larger max_cand
 * The first call, and any larger one, grows a plan-owned workspace (batch x
 * groups x n^3 float32 maps plus candidate lists).  With an MFMA x pass
 * selected, the float32 maps of the candidates come from the FFT x pass.
 */
int vx_rescore_f64(vx_plan* plan, const double* tdata_dev, const float* tvec_dev, const float* tprob_dev,
                   double t_ave, const double* sdata_dev, const float* svec_dev, const float* sprob_dev,
                   const float* rot_dev, int nrows, double window, int max_cand, double* score_dev,
                   int32_t* trans_dev, int32_t* ncand_dev, vx_stream stream);

/*
 * vx_score_rows with a HIP event recorded on `stream` between the launches of
 * every batch; the stream is synchronised after each batch.  stage_ms_host[6]
 * receives the summed milliseconds of {gather, forward z/y FFT, x pass with
 * conjugate multiply and channel sum, inverse y/z FFT with argmax, prob-mode
 * mix, finalize}.  For measurement; results equal vx_score_rows'.
 */
int vx_score_rows_profile(vx_plan* plan, const float* rot_dev, int nrows, float* best_dev,
                          int32_t* trans_dev, double* stage_ms_host, vx_stream stream);

/*
 * Top-N rows by (score descending, row ascending) -- the stable
 * sorted(..., reverse=True)[:TopN] of VESPER_1_prob_new.py:664, :739, :992,
 * :1082.  Scores are scores_dev[row*stride + col]; rows_dev receives
 * min(topn, nrows) row indices.  topn <= 1024.  The key buffers come from the
 * library's grow-on-demand scratch (see Conventions).
 */
int vx_topn(const float* scores_dev, int nrows, int stride, int col, int topn,
            int32_t* rows_dev, vx_stream stream);

/*
 * vx_topn on a caller-owned workspace (for HIP-graph capture and callers that
 * keep their own memory): ws_dev holds at least *bytes_host bytes as
 * vx_topn_ws_bytes gives them for (nrows, topn), 8-byte aligned.  Allocates
 * nothing; results equal vx_topn's.  Calls sharing one workspace must not
 * overlap.
 */
int vx_topn_ws_bytes(int nrows, int topn, int64_t* bytes_host);
int vx_topn_ws(const float* scores_dev, int nrows, int stride, int col, int topn,
               int32_t* rows_dev, void* ws_dev, int64_t ws_bytes, vx_stream stream);

/*
 * vx_topn that allocates nothing, for callers that cannot allocate device
 * memory (code under HIP-graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  No counterpart in the reference.  Its key buffers are a
 * 2.1 MB array in the library's code object that each GPU holds from load to
 * process end; nrows <= 262144 (2^18), else VX_ERR_ARG.  The first call on a
 * device resolves that array's address there (hipGetSymbolAddress), so make
 * one call before capturing it.  All calls on one GPU share the array: calls
 * that overlap in time on one GPU (other streams, other host threads) race.
 * Results equal vx_topn's.
 */
int vx_topn_noalloc(const float* scores_dev, int nrows, int stride, int col, int topn,
                    int32_t* rows_dev, vx_stream stream);

/*
 * One rotation of the query map: rot_mrc (VESPER_1_prob_new.py:279-354) or,
 * with prob_dev != NULL, rot_mrc_prob (:1211-1299).  rot_host9 = F as above.
 * Outputs are fully written (zeros where nothing lands).
 */
int vx_rotate_map(int n, const float* rot_host9, const double* data_dev,
                  const float* vec_dev, const float* prob_dev, double* out_data_dev,
                  float* out_vec_dev, float* out_prob_dev, int flags, vx_stream stream);

/*
 * The float32 query planes that vx_score_rows correlates for the rotation
 * rot_dev[9] (device): rot_mrc / rot_mrc_prob of the search map followed by
 * the mode's query channels (VESPER_1_prob_new.py:279-354, :1211-1299,
 * :546-558, :1434-1438, :1485-1497), formed by the search's own gather
 * kernel.  out_dev holds C planes [C][n][n][n] (C as vx_mode_planes),
 * zero outside the plan's support box.  Uses the plan's first workspace in
 * Laplacian mode: do not overlap it with a search on the same plan.
 */
int vx_query_planes(vx_plan* plan, const float* rot_dev, float* out_dev, vx_stream stream);

/*
 * Final FP64 re-scoring at a translation: get_score
 * (VESPER_1_prob_new.py:1302-1393), including its d2 = where(d2 <= 0, 0, d1)
 * line (:1357).  out_host[10] = {Nm, total, cc_raw, pcc_raw, dot_score,
 * ave2, std2, pstd2, overlap_count_target, search_positive_count}.
 */
int vx_get_score(int n, const double* tdata_dev, const float* tvec_dev, double t_ave,
                 const double* sdata_dev, const float* svec_dev,
                 const int32_t* trans3_host, double* out_host, vx_stream stream);

/*
 * Threshold and extent of a source map: mrc_set_vox_size
 * (VESPER_1_prob_new.py:78-85) sets data < th to 0 in place (th compared in
 * float32) and returns out2_host = {dmax, count of non-zero voxels}, dmax =
 * max over non-zero voxels of |index - dims/2| in voxel units.
 */
int vx_threshold_dmax(float* data_dev, int sx, int sy, int sz, float th, double* out2_host,
                      vx_stream stream);

/*
 * The float32 FFT input planes of a map for a mode: the vector components of
 * rot_and_search_fft / rot_and_search_fft_prob (VESPER_1_prob_new.py:1434-1438,
 * :1485-1497), the query transform of fft_search_score_trans_1d (:546-558:
 * Overlap 1[d>0], CC d[d>0], PCC (d - ave)[d>0], Laplacian laplacian_filter(d))
 * and the target channels (:601-615, :831-838).  Values are formed in float64,
 * then rounded to float32.  out_dev holds C planes [C][n][n][n]: C = 3 for
 * VX_MODE_VEC (vec_dev [n^3 x 3]), 1 for O/C/P/L (data_dev float64 [n^3]),
 * 7 for VX_MODE_PROB (3 vector + the 4 planes of prob_dev [4][n^3]).
 */
int vx_mode_planes(int n, int mode, const double* data_dev, const float* vec_dev, const float* prob_dev,
                   double ave, float* out_dev, vx_stream stream);

/* laplacian_filter (VESPER_1_prob_new.py:494-526) on an [n][n][n] float64 grid. */
int vx_laplacian(int n, const double* in_dev, double* out_dev, vx_stream stream);

/*
 * interpolate (interpolation_functions.py:6-135): fill the zero entries off
 * the reference parity lattice of a class-prediction grid [sx][sy][sz][nl]
 * (float32, or float64 with is_f64 = 1), in place on the device.  The
 * reference parity is that of the first voxel in C order whose channel 0 is
 * positive (returned in ref_host3 when non-NULL); a grid without one is an
 * error.  Seven passes as the reference's loops, with its index -1
 * wrap-around and float64 averages.  Synchronises the stream once.
 */
int vx_interpolate(void* data_dev, int is_f64, int sx, int sy, int sz, int nl, int32_t* ref_host3,
                   vx_stream stream);

/*
 * fft_search_best_dot (VESPER_1_prob_new.py:464-491) and the transform of
 * fft_search_score_trans_1d (:529-565): for each of nch channels,
 * out_dev[c] = irfftn(tspec_dev[c] * rfftn(q_dev[c])), unnormalised, with the
 * caller's conjugated target spectra tspec_dev (nch x [n][n][n/2+1]
 * complex64, the layout of numpy's rfftn), queries q_dev (nch x [n]^3
 * float32) and outputs out_dev (nch x [n]^3 float32).  Uses the plan's FFT
 * workspace (any mode; the plan's grid edge); does not touch its target or
 * query.
 */
int vx_correlate(vx_plan* plan, const void* tspec_dev, int nch, const float* q_dev, float* out_dev,
                 vx_stream stream);

/*
 * find_best_trans_list (VESPER_1_prob_new.py:357-374): the float32 channel
 * sum of nmaps (<= 8) real correlation maps of n3 voxels each, its maximum and
 * first (C-order) argmax.  maps_host holds nmaps device pointers.  Writes
 * best_dev[0] and index_dev[0] (flat index; the int64 doubles as scratch).
 */
int vx_sum_argmax(const float* const* maps_host, int nmaps, int64_t n3, float* best_dev,
                  int64_t* index_dev, vx_stream stream);

/*
 * find_best_trans_mixed (VESPER_1_prob_new.py:1530-1545): Python-sum of each
 * group, float32 z-scores (vave, vstd, pave, pstd rounded to float32) and the
 * (1 - alpha), alpha mix; maximum and first argmax as vx_sum_argmax.
 */
int vx_mixed_argmax(const float* const* vmaps_host, int nv, const float* const* pmaps_host, int np,
                    int64_t n3, double alpha, double vstd, double vave, double pstd, double pave,
                    float* best_dev, int64_t* index_dev, vx_stream stream);

/*
 * sco_arr of get_score (VESPER_1_prob_new.py:1362-1370): float32 per-voxel
 * products tvec . svec placed at the search positions t + trans (trans
 * wrapped as the reference does), zero elsewhere; out_dev is [n]^3 float64
 * (out_f64 = 1, the orig mode's dtype) or float32 (prob mode).
 */
int vx_sco_arr(int n, const float* tvec_dev, const float* svec_dev, const int32_t* trans_host3,
               int out_f64, void* out_dev, vx_stream stream);


/* ------------------------------------------------------------------------
 * Learned rotation proposal.  No counterpart in the reference (which scores
 * every Euler row, VESPER_1_prob_new.py:644-655): a 4-layer MLP predicts a
 * row's DOT score so that only the best fraction of rows is FFT-scored.
 * bf16 matrices are raw uint16 storage; fp32 accumulation on MFMA.
 * ------------------------------------------------------------------------ */

/* C = op(A) op(B): A [M x K] (transA: stored K x M), B [K x N] (transB:
 * stored N x K), row-major bf16.  K a multiple of 16.
 * epilogue 0: C fp32 = AB;  1: C bf16 = relu(AB + bias);  2: C bf16 =
 * AB where aux (bf16 [M x N], ld ldaux) > 0 else 0;  3: C fp32 += AB (one
 * workgroup per element of C, so the result does not depend on timing). */
int vx_gemm_bf16(int M, int N, int K, const void* A_dev, int lda, int transA, const void* B_dev, int ldb,
                 int transB, void* C_dev, int ldc, int epilogue, const float* bias_dev, const void* aux_dev,
                 int ldaux, vx_stream stream);

/* output layer pred[m] = sum_k X[m][k] w[k] + b[0] (X bf16 [M x K], ld ldx) */
int vx_mlp_head_fwd(int M, int K, const void* X_dev, int ldx, const float* w_dev, const float* b_dev,
                    float* pred_dev, vx_stream stream);

/* its backward with mean-squared loss: dw += .., db += .., dX = d w (X > 0),
 * loss += mean (pred - y)^2; the sums over rows run in a fixed order, so
 * repeated calls give identical results */
int vx_mlp_head_bwd(int M, int K, const void* X_dev, int ldx, const float* w_dev, const float* pred_dev,
                    const float* y_dev, float* dw_dev, float* db_dev, void* dX_dev, float* loss_dev,
                    vx_stream stream);

/* out[n] += sum_m G[m][n] (bf16 G), summed in a fixed order */
int vx_colsum_bf16(int M, int N, const void* G_dev, int ldg, float* out_dev, vx_stream stream);

/* Adam step (bias-corrected, step >= 1) on fp32 p; refreshes the bf16 copy if non-NULL */
int vx_adam(int n, float* p_dev, const float* g_dev, float* m1_dev, float* m2_dev, void* p_bf16_dev, float lr,
            float beta1, float beta2, float eps, int step, float weight_decay, vx_stream stream);

int vx_cast_bf16(int n, const float* in_dev, void* out_bf16_dev, vx_stream stream);

/* 32 rotational moments of a resampled map (data float64 [n^3], vec float32
 * [n^3 x 3]) about the rotation centre n/2: centroid, second moments, mean
 * vector, position-vector and vector-vector moments, 5 invariants. */
int vx_mlp_descriptors(int n, const double* data_dev, const float* vec_dev, float* out32_dev, vx_stream stream);

/* Feature rows (bf16 [M x ld], ld >= 68): quaternion, target moments, query
 * moments rotated by R = F^T of each row of rot_dev [M x 9]. */
int vx_mlp_features(int M, int ld, const float* rot_dev, const float* quat_dev, const float* tdesc_dev,
                    const float* qdesc_dev, void* X_dev, vx_stream stream);

/* Rows whose prediction ranks in about the top k (4096-bin histogram
 * threshold; ties at the threshold bin are all kept), in ascending order.
 * rows_dev [M]; *nsel_host = number written (>= k unless M < k).
 * Synchronises. */
int vx_select_top(int M, const float* pred_dev, int k, int* rows_dev, int* nsel_host, vx_stream stream);

/* Proposal bookkeeping (device-side; no counterpart in the reference). */
/* out = (in - mean) / (std + 1e-30), std with n - 1; stats_dev[2] = {mean,
 * std + 1e-30} computed in FP64 in a fixed order (use_given = 0) or read
 * (use_given = 1).  in and out may alias. */
int vx_zscore_f32(int n, const float* in_dev, double* stats_dev, int use_given, float* out_dev, vx_stream stream);
/* dst row r = src row idx[r] for r < nvalid, zero rows for nvalid <= r < nout;
 * rows of row_bytes bytes (a multiple of 4). */
int vx_gather_rows(int nout, int nvalid, const int32_t* idx_dev, const void* src_dev, int row_bytes, void* dst_dev,
                   vx_stream stream);
/* pred[i] for scored[i] != 0 set just below the lowest unscored prediction. */
int vx_mask_scored(int n, float* pred_dev, const uint8_t* scored_dev, vx_stream stream);
/* out = the rows[i] with scored[rows[i]] == 0, in order; *count_host their
 * number.  Synchronises. */
int vx_filter_unscored(int n, const int32_t* rows_dev, const uint8_t* scored_dev, int32_t* out_dev, int* count_host,
                       vx_stream stream);
/* dst[idx[i]] = src[i * stride]; flag[idx[i]] = 1 when flag_dev != NULL. */
int vx_scatter_scores(int n, const int32_t* idx_dev, const float* src_dev, int stride, float* dst_dev,
                      uint8_t* flag_dev, vx_stream stream);
int vx_fill_f32(int n, float value, float* dst_dev, vx_stream stream);

#ifdef __cplusplus
}
#endif

#endif /* VESPER_HIP_H */
