/*
 * v3dmatch.h -- C ABI of the MI355X SIFT putative-matching engine.
 *
 * The entry points replace the device side of V3D_CUDA::SIFT_Matching
 * (reference: V3D/include/CUDA/v3d_cudamatching.h:15-67,
 * V3D/src/CUDA/v3d_cudamatching.cpp:14-255) and its internal C ABI
 * (V3D/include/CUDA/v3d_cudamatching.cuh:26-32).  Plain pointers and sizes
 * only; every call returns a status code (V3DM_OK == 0) and the message of
 * the last failure is available from v3dm_last_error().
 *
 * Descriptors are 128-D, row-major, one row per feature.  Float descriptors
 * are what the reference receives; uint8 descriptors u are the same data as
 * the floats u/512 (the reference matcher sees exactly those floats).
 * Match lists are (left index, right index) int32 pairs in ascending left
 * order, as produced by the reference mutual check
 * (v3d_cudamatching.cpp:121-126).
 *
 * Threading: a context owns one HIP stream and its device buffers.  Calls on
 * one context must not overlap; several contexts may coexist.
 */
#ifndef V3DMATCH_H
#define V3DMATCH_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* status codes */
#define V3DM_OK 0
#define V3DM_ERR_ARG (-1)         /* invalid argument */
#define V3DM_ERR_STATE (-2)       /* call out of order (e.g. find before allocate) */
#define V3DM_ERR_DEVICE (-3)      /* HIP runtime error or no usable device */
#define V3DM_ERR_NOMEM (-4)       /* device or pinned-host allocation failed */
#define V3DM_ERR_UNSUPPORTED (-5) /* request outside what the engine implements */
#define V3DM_ERR_CAPACITY (-6)    /* caller's output buffer too small; *count holds the size needed */

/* kernel selection for v3dm_set_kernel() */
#define V3DM_KERNEL_AUTO 0  /* int8 MFMA when both sides are u8-representable, else BF16 + FP32 rescoring */
#define V3DM_KERNEL_PLAIN 1 /* reference design: FP32 score matrix in HBM + column/row passes */
#define V3DM_KERNEL_I8 2    /* fused int8 MFMA matcher (bit-exact for u8 / u-over-512 inputs) */
#define V3DM_KERNEL_BF16 3  /* float descriptors: BF16 MFMA candidates within a proven error bound of the
                               decision threshold, rescored with the plain kernel's FP32 arithmetic
                               (same match list as V3DM_KERNEL_PLAIN) */

typedef struct v3dm_ctx v3dm_ctx;

/* --- lifecycle ----------------------------------------------------------- */

/* SIFT_Matching::SIFT_Matching() -- v3d_cudamatching.h:17-21.  Binds the
 * context to HIP device `device` and creates its stream. */
int v3dm_create(int device, v3dm_ctx **out);

/* Frees everything the context holds (the reference leaves this to
 * deallocate(); destroying also deallocates). */
void v3dm_destroy(v3dm_ctx *ctx);

/* Message of the last failed call on ctx ("" if none); ctx may be NULL for
 * failures of v3dm_create. */
const char *v3dm_last_error(const v3dm_ctx *ctx);

/* enableRoundingWarning(bool) -- v3d_cudamatching.h:23-26 */
int v3dm_set_rounding_warning(v3dm_ctx *ctx, int flag);

/* enableReductionPath(bool) -- v3d_cudamatching.h:28-31.  Accepted and
 * recorded; the match lists are those of the default (non-reduction) path. */
int v3dm_set_reduction_path(v3dm_ctx *ctx, int flag);

/* allocate(nMaxLeft, nMaxRight, useGuidedMatching) --
 * v3d_cudamatching.h:33-37, v3d_cudamatching.cpp:14-49.  Maxima are rounded
 * down to a multiple of 32 (warning on stderr when enabled). */
int v3dm_allocate(v3dm_ctx *ctx, int nMaxLeft, int nMaxRight, int useGuidedMatching);

/* deallocate() -- v3d_cudamatching.cpp:51-63 */
int v3dm_deallocate(v3dm_ctx *ctx);

/* --- descriptors ---------------------------------------------------------- */

/* setLeftFeatures / setRightFeatures -- v3d_cudamatching.h:41-42,
 * v3d_cudamatching.cpp:65-95.  n is rounded down to a multiple of 32
 * (warning when enabled) and clamped to the allocated maximum; the host
 * array is copied before the call returns. */
int v3dm_set_left_f32(v3dm_ctx *ctx, int n, const float *desc);
int v3dm_set_right_f32(v3dm_ctx *ctx, int n, const float *desc);

/* Additive: the same calls for uint8 descriptors (equivalent to floats u/512). */
int v3dm_set_left_u8(v3dm_ctx *ctx, int n, const uint8_t *desc);
int v3dm_set_right_u8(v3dm_ctx *ctx, int n, const uint8_t *desc);

/* Current (rounded, clamped) feature counts. */
int v3dm_get_counts(const v3dm_ctx *ctx, int *nLeft, int *nRight);

/* Device buffers of the current state, for code that read SIFT_Matching's
 * protected _d_leftFeatures, _d_rightFeatures, _d_scores, _d_leftPositions and
 * _d_rightPositions (v3d_cudamatching.h:60-66).  Each output may be NULL, and
 * receives NULL when this engine holds no such buffer:
 *  - left / right: the FP32 descriptors (n x 128, row-major); present for float
 *    input, and for uint8 input once the plain path has run;
 *  - scores: the nRight x nLeft FP32 score matrix with guided masks applied
 *    (the reference's layout), present only when the plain path produced the
 *    last list (the int8 and BF16 scans never store scores);
 *  - leftPos / rightPos: the first guided call's positions since the last
 *    findPutativeMatches, interleaved (x, y) (the reference stores all x,
 *    then all y).
 * The pointers stay valid until the next call that changes the state. */
int v3dm_get_device_buffers(const v3dm_ctx *ctx, const float **left, const float **right, const float **scores,
                            const float **leftPos, const float **rightPos);

/* --- matching ------------------------------------------------------------- */

/* findPutativeMatches(matches, minScore, minRatio) --
 * v3d_cudamatching.h:44-46, v3d_cudamatching.cpp:97-127.  Ratio test off
 * when minRatio < 0.  Writes min(count, capacity) pairs (2 int32 each);
 * returns V3DM_ERR_CAPACITY when count > capacity. */
int v3dm_find_putative_matches(v3dm_ctx *ctx, float minScore, float minRatio, int32_t *pairs, int capacity,
                               int *count);

/* runGuidedMatching(F, distanceThreshold, leftPositions, rightPositions,
 * matches, minScore, minRatio) -- v3d_cudamatching.h:48-52,
 * v3d_cudamatching.cpp:129-191, v3d_cudamatching.cu:259-293.  F: row-major
 * 3x3 doubles; positions: interleaved (x, y) floats for the current left and
 * right features.  Like the reference, the mask accumulates with the masks
 * of earlier guided calls since the last findPutativeMatches.  Requires
 * allocate(..., useGuidedMatching = 1). */
int v3dm_run_guided_matching(v3dm_ctx *ctx, const double F[9], float distanceThreshold, const float *leftPos,
                             const float *rightPos, float minScore, float minRatio, int32_t *pairs, int capacity,
                             int *count);

/* runGuidedHomographyMatching(...) -- v3d_cudamatching.h:54-58,
 * v3d_cudamatching.cpp:193-255, v3d_cudamatching.cu:295-325. */
int v3dm_run_guided_homography_matching(v3dm_ctx *ctx, const double H[9], float distanceThreshold,
                                        const float *leftPos, const float *rightPos, float minScore,
                                        float minRatio, int32_t *pairs, int capacity, int *count);

/* Scan tile variant of the single-pair int8 path (0 .. v3dm_scan_variant_count()-1),
 * or -1 (default): the fastest, measured once per shape bucket (each side's
 * count rounded up to a power of two) on the first job of that bucket and
 * cached in the context.  No counterpart in the reference. */
int v3dm_set_scan_variant(v3dm_ctx *ctx, int variant);
/* The scan variant the last single-pair int8 job ran (-1 before any). */
int v3dm_get_last_scan_variant(const v3dm_ctx *ctx, int *variant);
/* Whether that job's scan applied its survivors itself (1: the whole grid was
 * resident at once, no separate insertion pass) or left them to the insertion
 * kernel (0); -1 before any job.  Diagnostic, no counterpart in the reference. */
int v3dm_get_last_scan_inline(const v3dm_ctx *ctx, int *inlineInsert);

/* Survivor insertion pass of int8 jobs whose scan grid exceeds what the
 * device holds at once: -1 (default) chooses per job (one wave per tile
 * above 128 survivor regions per CU, else one wave per region), 0 forces one
 * wave per region, 1 one wave per tile.  Both forms give the same tables.
 * Applies to single-pair, batched and learned-fallback jobs of the context.
 * No counterpart in the reference. */
int v3dm_set_insert_form(v3dm_ctx *ctx, int form);
/* Form the last single-pair or batched int8 job's insertion pass ran (0 or
 * 1), or -1 when its scan applied the survivors itself (or before any job). */
int v3dm_get_last_insert_form(const v3dm_ctx *ctx, int *form);

/* Kernel selection (V3DM_KERNEL_*) and the kernel the last match call ran. */
int v3dm_set_kernel(v3dm_ctx *ctx, int kernel);
int v3dm_get_last_kernel(const v3dm_ctx *ctx, int *kernel);

/* --- asynchronous form (device-resident inputs, no host round trip) ------- */

/* The context's HIP stream (hipStream_t) for event timing. */
int v3dm_get_stream(const v3dm_ctx *ctx, void **stream);

/* Enqueue findPutativeMatches on the context stream and return at once; the
 * per-left match table stays on the device until v3dm_fetch_matches(). */
int v3dm_find_async(v3dm_ctx *ctx, float minScore, float minRatio);

/* Wait for the last enqueued match and copy its list to the host. */
int v3dm_fetch_matches(v3dm_ctx *ctx, int32_t *pairs, int capacity, int *count);

int v3dm_synchronize(v3dm_ctx *ctx);

/* --- the same job with no device allocation ------------------------------ */

/* No counterpart in the reference: for callers that cannot allocate device
 * memory (code under HIP graph capture, tight serving loops, harnesses that
 * forbid hipMalloc).  findPutativeMatches of one uint8 pair whose
 * descriptors are already in device memory: dLeft (nLeft x 128) and dRight
 * (nRight x 128), row-major on the current HIP device; counts are rounded
 * down to multiples of 32 as setLeftFeatures does.  Writes the per-left
 * table (right index or -1, the reference's mutual list,
 * v3d_cudamatching.cpp:121-126) to dMatchOfLeft[0 .. rounded nLeft) on the
 * device.  Enqueued on `stream` (a hipStream_t, NULL = the null stream) and
 * returns at once; the caller synchronises.  The working memory is a
 * 3.4 MiB array of the library in each GPU's memory (held until the process
 * ends; the first call on a device looks up its address and clears it with
 * a kernel), so at most 8192 features per side (V3DM_ERR_CAPACITY beyond);
 * nothing is allocated by the call.  Every call on one GPU uses that GPU's
 * array: calls that overlap on one GPU (other streams, other threads) race.
 * Errors are reported by v3dm_last_error(NULL). */
int v3dm_find_u8_noalloc(const uint8_t *dLeft, int nLeft, const uint8_t *dRight, int nRight, float minScore,
                         float minRatio, int32_t *dMatchOfLeft, void *stream);

/* --- kernel timing (HIP events recorded on the context stream) ------------ */

/* When enabled, every match job records HIP events around its kernels
 * (scan: the fused int8 scan or, on the plain path, the score-matrix
 * kernel; select: finalize / selection passes).  Enabling resets the
 * totals.  Up to 65536 jobs are kept. */
int v3dm_profile_enable(v3dm_ctx *ctx, int enable);

/* Waits for the stream, then returns the summed event times in ms of the
 * jobs recorded since profiling was enabled, and their number. */
int v3dm_profile_read(v3dm_ctx *ctx, double *scanMs, double *selectMs, int *jobs);

/* --- many image pairs in one job -------------------------------------------- */
/* The reference matches one pair per findPutativeMatches call
 * (SfM/match_sequence/src/match_sequence.cpp:157-203,
 * SfM/match_using_voctree/src/match_using_voctree.cpp:277-323); these entry
 * points run a whole list of pairs in one job with the same per-pair rule,
 * so each pair's table equals findPutativeMatches on that pair. */

/* Uploads nImages uint8 descriptor sets (rows concatenated in image order,
 * counts[i] rows each).  Counts are rounded down to multiples of 32 like
 * setLeftFeatures/setRightFeatures (v3d_cudamatching.cpp:68-77).  Clears any
 * pair list. */
int v3dm_batch_set_images_u8(v3dm_ctx *ctx, int nImages, const int32_t *counts, const uint8_t *desc);

/* As v3dm_batch_set_images_u8 for float descriptors (the reference's
 * setLeftFeatures/setRightFeatures(int, const float*) input, e.g. SiftGPU's
 * unit descriptors): the batch then runs the BF16 candidate scan + FP32
 * rescoring, whose per-pair tables equal the plain FP32 rule's. */
int v3dm_batch_set_images_f32(v3dm_ctx *ctx, int nImages, const int32_t *counts, const float *desc);

/* Sets the pairs to match: pairImages[2p] = left image, [2p+1] = right image.
 * Clears the per-pair guided masks. */
int v3dm_batch_set_pairs(v3dm_ctx *ctx, int nPairs, const int32_t *pairImages);

/* Feature positions of the uploaded images: interleaved (x, y) floats per
 * feature, images in the order and with the counts of the last
 * v3dm_batch_set_images_* call -- the leftPositions / rightPositions that
 * runGuidedMatching takes (v3d_cudamatching.h:48-52; the drivers pass each
 * view's SIFT_Feature::position, match_sequence.cpp:285-292).  Needed by
 * v3dm_batch_set_guides; cleared by v3dm_batch_set_images_*. */
int v3dm_batch_set_positions(v3dm_ctx *ctx, const float *positions);

/* Guided matching for every pair of the current pair list in one job: for
 * pair p, kind[p] = 0 leaves the pair's findPutativeMatches list, 1 gives
 * runGuidedMatching(F, distanceThreshold[p], ...) with the row-major
 * fundamental matrix F = M[9p .. 9p+8] (v3d_cudamatching.cpp:129-191,
 * v3d_cudamatching.cu:259-293), 2 gives runGuidedHomographyMatching with the
 * homography H = M[9p .. 9p+8] (.cpp:193-255, .cu:295-325).  Each pair's
 * table then equals findPutativeMatches followed by that one guided call on
 * the pair -- what the drivers do after robust orientation
 * (match_sequence.cpp:275-293).  The masks stay for later jobs until the next
 * v3dm_batch_set_pairs / set_images or a call with kind = NULL (clears them).
 * Requires v3dm_batch_set_positions.  No counterpart in the reference, which
 * runs one guided call per pair. */
int v3dm_batch_set_guides(v3dm_ctx *ctx, const int32_t *kind, const double *M, const float *distanceThreshold);

/* Runs the job.  matchOfLeft receives, for pair p, entries
 * [offsets[p], offsets[p+1]) = per-left-feature right index or -1 (the
 * reference's mutual list as a table); offsets has nPairs + 1 entries. */
int v3dm_batch_match(v3dm_ctx *ctx, float minScore, float minRatio, int32_t *matchOfLeft, int64_t *offsets);

/* Asynchronous form: enqueue, then fetch (waits). */
int v3dm_batch_match_async(v3dm_ctx *ctx, float minScore, float minRatio);
int v3dm_batch_fetch(v3dm_ctx *ctx, int32_t *matchOfLeft, int64_t *offsets);

/* Scan tile variant: -1 = automatic (tuning cache named by the environment
 * variable V3DM_TUNE_CACHE, else a timed trial of every variant on the
 * first job); otherwise a variant index. */
int v3dm_batch_set_variant(v3dm_ctx *ctx, int variant);
int v3dm_batch_get_variant(const v3dm_ctx *ctx, int *variant);

/* Times every variant on the current job (reps runs each), keeps the fastest;
 * msPerVariant (may be NULL) receives v3dm_scan_variant_count() times. */
int v3dm_batch_autotune(v3dm_ctx *ctx, float minScore, float minRatio, int reps, double *msPerVariant, int *best);

/* Capture the job's kernels in a hipGraph and replay it while the pair set
 * and thresholds stay the same. */
int v3dm_batch_set_graph(v3dm_ctx *ctx, int enable);

/* Pair count, total left features, scan tiles of the current variant, and
 * descriptor pairs scored (sum of nL * nR) of the current job. */
int v3dm_batch_info(const v3dm_ctx *ctx, int *nPairs, int64_t *nLeftTotal, int *nTiles, int64_t *pairsScored);

/* Scan tile variants: count, and waves per workgroup / tile rows / tile columns. */
int v3dm_scan_variant_count(void);
/* Variants 0 .. v3dm_scan_variant_count_f32()-1 also have a float-descriptor
 * (BF16) form; a float batch accepts only those. */
int v3dm_scan_variant_count_f32(void);
int v3dm_scan_variant_info(int variant, int *waves, int *tileRows, int *tileCols);
/* *wholeTile = 1 when the variant stages its whole tile (both sides' rows) in
 * LDS at the tile's start and runs its steps with no barrier between them
 * (kernel k_i8_scan_whole; one workgroup per CU), 0 for the ring form
 * (k_i8_scan: right rows through a four-group LDS ring, left rows in
 * registers). */
int v3dm_scan_variant_form(int variant, int *wholeTile);

/* --- synthetic SIFT-like descriptors (host) -------------------------------- */

/* Generates one image pair: nL left descriptors, nR right descriptors of
 * which min(floor(rho*nR), nL) are noisy copies (N(0, sigma) per component
 * before renormalisation) of distinct left rows at random positions.  Any
 * output pointer may be NULL.  rightSource[r] = source left row or -1. */
int v3dm_synth_pair(uint64_t seed, int nL, int nR, double rho, double sigma, int nThreads, uint8_t *leftU8,
                    uint8_t *rightU8, float *leftF32, float *rightF32, int32_t *rightSource);

/* Seed of pair `pairIndex` of workload `config`: 0x5EED0000 + 256*config + pairIndex. */
uint64_t v3dm_synth_seed(int config, int pairIndex);

/* --- diagnostics ----------------------------------------------------------- */

/* Number of HIP devices visible (0 without a GPU). */
int v3dm_device_count(int *count);

/* Runs one v_mfma_i32_32x32x32_i8 on device `device`, operands packed the way
 * the matcher packs descriptors: row i of A holds (i, 32) in its first two
 * bytes, column j of B holds (1, j), all other bytes 0, so D[i][j] = i + 32 j.
 * Writes the 64 lanes x 16 accumulator registers (lane-major, 1024 int32) to
 * out.  Tests use it to check the accumulator layout the matcher assumes:
 * register v of lane l holds row 8*(v/4) + 4*(l/32) + v%4, column l%32. */
int v3dm_mfma_layout_probe(int device, int32_t *out);

/* Rate of the int8 scan's step loop alone on given data: 8
 * v_mfma_i32_32x32x32_i8 per step on 64 rows per wave held in registers, the
 * max / compare test, the column term and (fromLds != 0) the four right
 * fragments read from LDS in the scan's swizzled layout, two waves per SIMD,
 * no global loads, barriers or survivors.  rows: 128 rows x 128 bytes in the
 * scan's stored form (u ^ 0x80), used as the left rows and the LDS stages.
 * The int8 MFMA's clock under its power limit depends on the operand values,
 * so this is the practical ceiling of the scan's inner loop on that data.
 * Writes the rate in TOP/s (one pair = 256 ops) to *topsOut.  Allocates and
 * synchronises: a diagnostic, not for use beside running jobs. */
int v3dm_mfma_ceiling_probe(int device, const uint8_t *rows, int fromLds, int iters, double *topsOut);

/* ---- learned proposal path (no counterpart in the reference) -------------
 *
 * An embedding network (128 -> 256 GELU -> 64, L2-normalised, BF16 weights,
 * input u / 64) trained on the GPU on matched descriptor pairs with a
 * symmetric InfoNCE loss, and a matcher that scores only each left's top-16
 * rights in the embedding space (exact int8 scores), rescans every left or
 * right query whose best - second margin is under `margin` exhaustively, and
 * applies the reference rule (v3d_cudamatching.cu:117-194,
 * v3d_cudamatching.cpp:121-126) to the exact scores of the scored pairs.
 * Parameter layout: W1 (256 x 128, row-major) | b1 (256) | W2 (64 x 256) | b2 (64).
 * A network object is not thread-safe. */
typedef struct v3dm_mlp v3dm_mlp;
int v3dm_mlp_param_count(void);
int v3dm_mlp_create(int device, uint64_t seed, v3dm_mlp **out);
void v3dm_mlp_destroy(v3dm_mlp *net);
const char *v3dm_mlp_last_error(const v3dm_mlp *net);
int v3dm_mlp_get_params(v3dm_mlp *net, float *params);
/* Sets the FP32 master parameters and resets the optimiser state. */
int v3dm_mlp_set_params(v3dm_mlp *net, const float *params);
/* Saves the network as inference uses it: an 8-byte tag "V3DMMLP\0", six
 * little-endian uint32 (version 1, widths 128 / 256 / 64, parameter count,
 * FNV-1a checksum of the payload), then W1 and W2 as BF16 and b1, b2 as FP32
 * in the parameter layout (99584 bytes).  v3dm_mlp_load checks all of it and
 * sets the masters to exactly those values (a loaded network embeds bit for
 * bit like the saved one; optimiser state reset).  No counterpart in the
 * reference. */
int v3dm_mlp_save(v3dm_mlp *net, const char *path);
int v3dm_mlp_load(v3dm_mlp *net, const char *path);
/* Loss and parameter gradient of one batch of matched pairs (left[i], right[i]),
 * batch x 128 uint8 each, at the given temperature; no update. */
int v3dm_mlp_loss_grad(v3dm_mlp *net, int batch, const uint8_t *left, const uint8_t *right, float temperature,
                       float *loss, float *grad);
/* Uploads the training set: n matched pairs, n x 128 uint8 per side. */
int v3dm_mlp_set_training_pairs(v3dm_mlp *net, int64_t n, const uint8_t *left, const uint8_t *right);
/* `steps` Adam steps (betas 0.9 / 0.999), each on `batch` pairs drawn with
 * replacement (seeded); losses[steps] receives the batch losses (may be NULL). */
int v3dm_mlp_train(v3dm_mlp *net, int steps, int batch, float lr, float temperature, uint64_t seed, float *losses);
/* Embeddings (n x 64 floats) of n uint8 descriptors. */
int v3dm_mlp_embed(v3dm_mlp *net, int n, const uint8_t *desc, float *out);
/* Learned matching of the context's current pair (both sides uploaded as
 * uint8), topK = 16 or 32 candidates per left.  Output as
 * v3dm_find_putative_matches; stats (may be NULL) receives [nLeft, nRight,
 * candidate pairs, flagged lefts, flagged rights, pairs scored by the
 * fallback scans]. */
int v3dm_learned_find(v3dm_ctx *ctx, v3dm_mlp *net, int topK, float margin, float minScore, float minRatio,
                      int32_t *pairs, int capacity, int *count, int64_t *stats);
/* Diagnostic: the last learned job's candidate table, n x topK right indices
 * (-1 for none), n <= the left count of that job; *topK receives K. */
int v3dm_learned_candidates(v3dm_ctx *ctx, int n, int32_t *cand, int *topK);

/* Diagnostic, no counterpart in the reference: per left feature k < n of the
 * last completed match job, the best and second-best score among the
 * candidates that reached the exact top-2 (scores above the survivor
 * threshold; -1e30 where there is none) and the right index of the best (-1
 * where none).  Exact integer-derived FP32 scores on the int8 path, rescored
 * FP32 scores on the BF16 path; not available after a plain-kernel job. */
int v3dm_fetch_left_top2(v3dm_ctx *ctx, int n, float *best, float *second, int32_t *argBest);

/* Diagnostic, no counterpart in the reference: copies `bytes` (a multiple of
 * 16) between two fresh device buffers `reps` times with a plain coalesced
 * copy kernel (k_copy_probe), each launch reading and writing exactly `bytes`.
 * Memory-counter passes (FETCH_SIZE, WRITE_SIZE) run it beside the matcher to
 * scale their readings to bytes. */
int v3dm_copy_probe(int device, int64_t bytes, int reps);

#ifdef __cplusplus
}
#endif

#endif /* V3DMATCH_H */
