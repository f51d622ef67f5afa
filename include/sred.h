/*
 * sred.h -- C ABI of libsred_hip.so, the MI355X (gfx950) implementation of
 * SReD's 2D block-correlation path.
 *
 * The library replaces the device side of SReD's CLUtils (JOCL/OpenCL) for
 * 2D global and block repetition.  Every entry point is extern "C", takes
 * plain pointers and sizes, returns an int status (SRED_OK == 0, < 0 on
 * error) and never throws; sred_last_error() gives the message of the last
 * failure on the calling thread.
 *
 * Buffers: entry points without a _dev suffix take HOST pointers (what a
 * JNI binding of CLUtils passes: Java float[] arrays) and are synchronous.
 * _dev entry points take DEVICE pointers and enqueue on the context's stream
 * (sred_ctx_set_stream); they return when the work is enqueued, except where
 * stated.  All images are row-major float32, index y*width+x.
 *
 * Citations name the reference (SReD) file:line each entry point replaces.
 */
#ifndef SRED_H
#define SRED_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- status codes ---- */
#define SRED_OK 0
#define SRED_ERR_ARG (-1)            /* invalid argument (even block size, NULL, size out of range) */
#define SRED_ERR_HIP (-2)            /* HIP runtime failure */
#define SRED_ERR_NONBINARY_MASK (-3) /* Utils.java:1002-1004 IllegalArgumentException */
#define SRED_ERR_UNKNOWN_METRIC (-4) /* CLUtils.java:2193-2195 returns null */
#define SRED_ERR_NONFINITE (-5)      /* Utils.java:861-884 NaN/Inf in a noise tile */
#define SRED_ERR_NO_DEVICE (-6)      /* no HIP device visible */

/* ---- metrics (GlobalRepetition2D_.java:20-25, BlockRepetition2D_.java:18-23) ---- */
#define SRED_PEARSON 0
#define SRED_SSIM 1
#define SRED_NRMSE 2
#define SRED_ABSDIFF 3

/* ---- contraction precision of the global maps ---- */
#define SRED_FP32 0   /* plain FP32 vector-ALU contraction */
#define SRED_BF16X3 1 /* split-BF16 (hi*hi + hi*lo + lo*hi) on MFMA, FP32 accumulate; blocks of more than
                         384 pixels (25x25 and larger; 385 is reflected into 384 dimensions) take the FP32
                         contraction, so every odd block size the reference accepts works in either
                         precision (the non-symmetric comparison contraction, SRED_OPT_SYMMETRIC = 0,
                         stops at 256 pixels) */

/* Largest supported block side (pixels); K <= 3200. */
#define SRED_MAX_BLOCK_SIDE 63

typedef struct sred_ctx sred_ctx;

/* ---- library / context ---- */
int sred_version(void);
const char *sred_last_error(void);
/* Replaces CLUtils.getOpenCLResources (CLUtils.java:169-238): one context per
 * device, reusable (unlike OpenCLResources, which each global call releases,
 * CLUtils.java:1773).  Not thread-safe: one context per caller thread. */
int sred_ctx_create(int device, sred_ctx **out);
int sred_ctx_destroy(sred_ctx *ctx);
/* Stream for _dev entry points (a hipStream_t; NULL = the null stream). */
int sred_ctx_set_stream(sred_ctx *ctx, void *stream);
int sred_ctx_synchronize(sred_ctx *ctx);
/* Column-chunk override for the global contraction (0 = automatic). */
int sred_ctx_set_chunk_cols(sred_ctx *ctx, int cols);
/* Tuning options of the global contraction (no counterpart in the
 * reference): SRED_OPT_CHUNK_COLS (as above); SRED_OPT_SYMMETRIC (1, the
 * default: the split-BF16 contraction evaluates each unordered block pair
 * once and adds it to both rows; 0: every ordered pair); for batched calls
 * SRED_OPT_AUTOTUNE (1, the default: time the panel sizes of a new class;
 * 0: the largest that fits), SRED_OPT_BATCH_PANEL (force a panel size: a
 * multiple of 256 up to 2048; 0 = off) and SRED_OPT_BATCH_MEMORY_MB (partial
 * sums per wave, default 16384); SRED_OPT_FIT_TIMING (0, the default; 1: a GAT
 * fit brackets every objective launch's row kernel with HIP events, summed
 * for sred_ctx_last_contraction, at about 4 us per launch);
 * SRED_OPT_PARTIAL_MEMORY_MB (partial sums of one global call's symmetric
 * contraction, default 16384: beyond it, from about 3e6 blocks (1730^2
 * pixels) on, the call runs the blocks as super-panels, one launch per
 * super-panel pair, with FP64 row totals in between, so any image size
 * fits). */
#define SRED_OPT_CHUNK_COLS 1
#define SRED_OPT_SYMMETRIC 2
#define SRED_OPT_AUTOTUNE 3
#define SRED_OPT_BATCH_PANEL 4
#define SRED_OPT_BATCH_MEMORY_MB 5
#define SRED_OPT_FIT_TIMING 6
#define SRED_OPT_PARTIAL_MEMORY_MB 7
int sred_ctx_set_option(sred_ctx *ctx, int option, int value);

/* Metric from the reference's dialog string; block != 0 uses the block
 * command's list ("NRMSE (inverted)", "Abs. Diff. of StdDevs.").
 * Mirrors the string dispatch of CLUtils.java:2174-2196 and :995-1010. */
int sred_metric_from_name(const char *name, int block);

/* ---- host-side mirrors of Utils (plain C++ on the host, no GPU) ---- */
/* Elliptical block size K: GlobalRepetition2D_.java:102-111. */
int sred_block_size_2d(int bw, int bh);
/* Min-max input normalisation: Utils.getInputImage2D, Utils.java:708-722. */
int sred_normalize_input_2d(float *img, int64_t n);
/* Utils.getReferenceBlock2D (Utils.java:557-639): px receives K pixels,
 * returns K (>0) or SRED_ERR_ARG for even sizes. */
int sred_reference_block_2d(const float *block, int bw, int bh, float *px, float *mean, float *std_);
/* Utils.getMeanNoiseVar2D (Utils.java:790-837). */
int sred_noise_var_2d(const float *img, int w, int h, float *noise_var);
/* Utils.getRelevanceMask (Utils.java:944-968) + nPixels
 * (GlobalRepetition2D_.java:133-140). */
int sred_relevance_mask_2d(const float *stds, int w, int h, int brw, int brh, float noise_var, float c,
                           float *mask, float *threshold, float *n_pixels);
/* Utils.applyMask2D (Utils.java:915-925), in place. */
int sred_apply_mask_2d(float *map, int w, int h, const float *mask);
/* Utils.normalizeImage2D (Utils.java:985-1031), in place; mask may be NULL. */
int sred_normalize_map_2d(float *map, int w, int h, int brw, int brh, const float *mask);

/* ---- device path, host buffers (the CLUtils methods a JNI binding replaces) ---- */
/* CLUtils.getLocalStatistics2D (CLUtils.java:376-467) with
 * kernelGetLocalStatistics2D.cl: float32 block mean and unbiased std. */
int sred_local_stats_2d(sred_ctx *ctx, const float *img, int w, int h, int bw, int bh, float *means,
                        float *stds);
/* Raw accumulators of kernelGetGlobal{Pearson,Ssim,Nrmse,AbsDiffStds}2D.cl
 * over all 64x64 launch tiles of CLUtils.java:1727-1740: S (similarity sum)
 * and Wsum (weight sum) per pixel.  filter_param and threshold are the values
 * the reference renders into the FILTER_PARAM/THRESHOLD macros
 * (noise variance, noise variance * relevance constant). */
int sred_global_accum_2d(sred_ctx *ctx, int metric, int precision, const float *img, const float *means,
                         const float *stds, int w, int h, int bw, int bh, float filter_param, float threshold,
                         float *S, float *Wsum);
/* CLUtils.calculateGlobalRepetitionMap2D (CLUtils.java:2174-2196) ->
 * getGlobal{Pearson,Ssim,Nrmse,AbsDiffStds}2D: accumulators plus the host
 * finalize (division by Wsum*nPixels+eps, inversion for NRMSE/AbsDiff, mask
 * if relevance_constant > 0, masked normalisation).  The inputs are those of
 * the Java call: the image, its local statistics, the relevance mask with its
 * constant, threshold and noise variance, and nPixels. */
int sred_global_repetition_2d(sred_ctx *ctx, int metric, int precision, const float *img, const float *means,
                              const float *stds, int w, int h, int bw, int bh, const float *mask,
                              float relevance_constant, float threshold, float noise_var, float n_pixels,
                              int normalize, float *out);
/* CLUtils.calculateBlockRepetitionMap2D (CLUtils.java:995-1010) ->
 * getBlock{Pearson,Ssim,Nrmse,AbsDiffStds}2D (CLUtils.java:486-981): local
 * statistics, template comparison, relevance mask and normalisation.
 * tmpl holds the K template pixels (ReferenceBlock2D.getPixels()). */
int sred_block_repetition_2d(sred_ctx *ctx, int metric, const float *img, int w, int h, const float *tmpl,
                             int bw, int bh, float tmpl_mean, float tmpl_std, float relevance_constant,
                             int normalize, float *out);

/* ---- device path, device buffers ---- */
/* Utils.getMeanNoiseVar2D (Utils.java:790-837) of a device image: the tile
 * variances (getMeanAndVarBlock2D, Utils.java:854-900, float sums in the
 * reference's order) on the GPU, the sort and the mean of the lowest
 * (int)(0.03 nBlocks + 1) on the host, as the _dev pipelines use it; the
 * same float as sred_noise_var_2d.  Synchronises. */
int sred_noise_var_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, float *noise_var);
/* Whole global-repetition pipeline of GlobalRepetition2D_.run
 * (GlobalRepetition2D_.java:124-145, input already normalised, VST off):
 * statistics, noise variance, relevance mask, nPixels, accumulators and
 * finalize.  d_img and d_out are device pointers (w*h floats).  Optional
 * device outputs (NULL to skip): d_means, d_stds, d_mask, d_S, d_Wsum.
 * Synchronises once on the stream (to read the 4-byte noise variance). */
int sred_global_map_2d_dev(sred_ctx *ctx, int metric, int precision, const float *d_img, int w, int h, int bw,
                           int bh, float relevance_constant, int normalize, float *d_out, float *d_means,
                           float *d_stds, float *d_mask, float *d_S, float *d_Wsum);
/* Whole block-repetition pipeline (BlockRepetition2D_.java:141-145 ->
 * CLUtils.java:995-1010) on device buffers; tmpl is a HOST array of K floats.
 * Synchronises once when relevance_constant != 0 (noise variance). */
int sred_block_map_2d_dev(sred_ctx *ctx, int metric, const float *d_img, int w, int h, const float *tmpl,
                          int bw, int bh, float tmpl_mean, float tmpl_std, float relevance_constant,
                          int normalize, float *d_out);
/* sred_block_map_2d_dev without device allocation.  No counterpart in the
 * reference: for callers that cannot allocate device memory (tight serving
 * loops over frames, harnesses that forbid hipMalloc).  Same arguments and
 * result as sred_block_map_2d_dev; every scratch buffer (template table,
 * min/max partials, normalisation words and, for relevance_constant != 0 or
 * NRMSE, the window statistics, mask and noise tiles) lives in one 64 MiB
 * device array per GPU that the library holds until the process ends.
 * Limit: calls whose scratch exceeds it are refused with SRED_ERR_ARG
 * (relevance 0 Pearson/SSIM/AbsDiff need 12 B per 1024 pixels beyond the
 * template; a relevance constant needs 12 B per pixel, up to about 5.5 M
 * pixels; NRMSE at relevance 0 needs 4 B per pixel, up to about 16 M).  Calls that overlap on one GPU share the array
 * and race; each call rewrites what it reads. */
int sred_block_map_2d_dev_noalloc(sred_ctx *ctx, int metric, const float *d_img, int w, int h, const float *tmpl,
                                  int bw, int bh, float tmpl_mean, float tmpl_std, float relevance_constant,
                                  int normalize, float *d_out);
/* The time-lapse branch of BlockRepetition2D_.run (BlockRepetition2D_.java:95-137)
 * on a stack of depth >= 3 frames (frame t at stack + t*w*h), without noise
 * stabilisation: the stack min-max normalised as a whole (getInputImage3D),
 * each frame normalised again (getInputImage2D(float[], ...)) and mapped as
 * sred_block_map_2d_dev does, then, when normalize, the stacked maps
 * normalised as a whole.  Synchronises once per frame when
 * relevance_constant != 0.  The _dev form takes device buffers (the template
 * stays a host array); the other copies the stack in and the maps out. */
int sred_block_timelapse_2d(sred_ctx *ctx, int metric, const float *stack, int w, int h, int depth,
                            const float *tmpl, int bw, int bh, float tmpl_mean, float tmpl_std,
                            float relevance_constant, int normalize, float *out);
int sred_block_timelapse_2d_dev(sred_ctx *ctx, int metric, const float *d_stack, int w, int h, int depth,
                                const float *tmpl, int bw, int bh, float tmpl_mean, float tmpl_std,
                                float relevance_constant, int normalize, float *d_out);

/* ---- batched global repetition (no counterpart in the reference: SReD
 * runs one image per command and loops stacks in Java) ---- */
typedef struct sred_global_problem {
    const float *d_img; /* device, w*h floats, input already normalised */
    int w, h, bw, bh;
    float *d_out;       /* device, w*h floats: the map of sred_global_map_2d_dev */
} sred_global_problem;
/* The maps of n images of any sizes and block sizes, each as
 * sred_global_map_2d_dev computes it with the split-BF16 contraction (same
 * relevance constant and normalisation for all).  The front steps run per
 * image; the contractions of all images run as persistent kernels (one per
 * block-size class and memory wave) inside ONE HIP graph launch together with
 * every image's reduction, finalize and normalisation.  Panel sizes come from
 * an autotuner (timed on the first batch holding a (metric, K_pad) class and
 * cached in the context) unless SRED_OPT_BATCH_PANEL forces one.  Metrics:
 * Pearson, SSIM, NRMSE (K <= 385).  Synchronises. */
int sred_global_batch_2d(sred_ctx *ctx, int metric, int n, const sred_global_problem *problems,
                         float relevance_constant, int normalize);
/* Panel size the autotuner chose on this context for (metric, k_steps =
 * ceil(K/32)), 0 if none yet. */
int sred_ctx_tuned_panel(sred_ctx *ctx, int metric, int k_steps, int *panel);
/* The autotuner's timed launch (ms, HIP events) for each candidate panel size
 * 256, 512, 1024, 2048 of that class on this context, 0 for candidates not
 * timed (more partial-sum memory than allowed, a forced or untuned panel). */
int sred_ctx_tuning_ms(sred_ctx *ctx, int metric, int k_steps, float *ms);

/* ---- Hu template metric (kernelGetPatchHu2D.cl:11-282; the reference has
 * the kernel but no Java caller, so these entries define its host side) ---- */
/* Map of sqrt(hu1(template) * hu1(window)) over every interior window, FP64
 * moments as the kernel computes them (window: mean over the square minus its
 * corners / K, min-max over the float ellipse, 1-based raw and 0-based central
 * moments).  tmpl is the bw x bh template RECTANGLE (row-major, host), used as
 * given inside the ellipse (the kernel's patch_pixels); border pixels are 0. */
int sred_block_hu_2d(sred_ctx *ctx, const float *img, int w, int h, const float *tmpl, int bw, int bh, float *out);
/* Same on device buffers (tmpl stays a host array). */
int sred_block_hu_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, const float *tmpl, int bw, int bh,
                         float *d_out);
/* Template preparation for the Hu map (the kernel expects a mean-subtracted,
 * normalised template and nothing in the reference prepares one): the
 * window's own steps (kernelGetPatchHu2D.cl:127-182) applied to the bw x bh
 * block rectangle; cells outside the ellipse are 0.  Host only. */
int sred_hu_prepare_template_2d(const float *blk, int bw, int bh, float *tmpl_out);

/* ---- variance stabilisation (the optional step before the path) ---- */
/* VarianceStabilisingTransform2D_.getGAT (VarianceStabilisingTransform2D_.java:136-151),
 * in place: v <- (g v + ref < 0) ? 0 : (2/g) sqrt(g v + ref),
 * ref = 3/8 g^2 + s^2 - g o, double arithmetic rounded to float (bit-exact).
 * Utils.getInputImage2D (Utils.java:697-702) passes the fitted parameters
 * rounded to float.  Synchronous; _dev enqueues on the context's stream. */
int sred_gat_2d(sred_ctx *ctx, float *img, int64_t n, double gain, double sigma, double offset);
int sred_gat_2d_dev(sred_ctx *ctx, float *d_img, int64_t n, double gain, double sigma, double offset);
/* GATMinimizer2D.userFunction (GATMinimizer2D.java:64-110): mean squared
 * deviation from 1 of the variances of the (w/6) x (h/6)-pixel blocks after
 * the fit's transform (special case v <= -ref/g).  Each block row's pixels,
 * each block's row sums and the block terms are summed in double over 64
 * lanes (element i to lane i mod 64, in order) combined by a fixed xor
 * butterfly.  Needs w, h >= 6. */
int sred_gat_objective_2d(sred_ctx *ctx, const float *img, int w, int h, double gain, double sigma, double offset,
                          double *error);
/* GATMinimizer2D.run (GATMinimizer2D.java:33-61) with the library's own
 * Nelder-Mead simplex in place of ImageJ's Minimizer (coefficients 1, 2,
 * 1/2, 1/2; initial simplex (gain0, sigma0, offset0) + (1, 10, 100) steps;
 * max_iter iterations or until every vertex has the same objective).
 * params[3] = (gain, sigma, offset); *error (may be NULL) = the objective
 * there.  The reference calls it with (1, 10, 100) and 5000 iterations. */
int sred_gat_minimize_2d(sred_ctx *ctx, const float *img, int w, int h, double gain0, double sigma0, double offset0,
                         int max_iter, double *params, double *error);
int sred_gat_minimize_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, double gain0, double sigma0,
                             double offset0, int max_iter, double *params, double *error);

/* The "Quad/Octree" estimate of Utils.getInputImage2D (Utils.java:686-703):
 * QuadTree_(w, h, min_size, alpha).buildTree, calculateRobustMeans(max_iter),
 * calculateLTSVariances(lts_fraction), collectMeanVariancePairs and
 * performLinearRegression (QuadTree_.java:135-407; the reference passes 4,
 * 0.01f, 50, 0.75f).  The node statistics, robust means and LTS sums keep
 * the reference's float arithmetic and sequential orders.  g0_edc[2] =
 * (g0, eDC); *n_leaves (may be NULL) = the leaf count; up to cap leaves are
 * written in pre-order to rects (x, y, w, h), means and variances (each may
 * be NULL).  Errors: SRED_ERR_NONFINITE for a NaN/Inf pixel or a NaN node
 * variance (Utils.java:879-895); SRED_ERR_ARG for a one-pixel image,
 * min_size < 1, alpha or lts_fraction outside [0, 1], or fewer than 3 leaves
 * (SimpleRegression.regress).  Synchronises. */
int sred_quadtree_gat_2d(sred_ctx *ctx, const float *img, int w, int h, int min_size, float alpha, int max_iter,
                         float lts_fraction, double *g0_edc, int *n_leaves, int cap, int32_t *rects, float *means,
                         float *variances);
int sred_quadtree_gat_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, int min_size, float alpha,
                             int max_iter, float lts_fraction, double *g0_edc, int *n_leaves, int cap,
                             int32_t *rects, float *means, float *variances);
/* QuadTree_.applyGATtree (QuadTree_.java:409-425) in place: ref = 3/8 g0^2 +
 * eDC in double, v <- (g0 v + ref < 0) ? 0 : (2/g0) sqrt(g0 v + ref), rounded
 * to float.  Synchronous; _dev enqueues on the context's stream. */
int sred_gat_tree_2d(sred_ctx *ctx, float *img, int64_t n, double g0, double edc);
int sred_gat_tree_2d_dev(sred_ctx *ctx, float *d_img, int64_t n, double g0, double edc);
/* Host-only helpers of the estimate (no GPU): the inverse CDF of the F(d1,
 * d2) distribution (commons-math3 FDistribution.inverseCumulativeProbability,
 * used at QuadTree_.java:241-245; bisection to adjacent doubles) and the
 * least-squares line of commons-math3 SimpleRegression (QuadTree_.java:395-407). */
int sred_f_quantile(double d1, double d2, double p, double *out);

/* ---- volumes (GlobalRepetition3D_) ---- */
/* Ellipsoid block size K of GlobalRepetition3D_.java:101-112 (odd sides,
 * depth >= 3), host only. */
int sred_block_size_3d(int bw, int bh, int bd);
/* CLUtils.getLocalStatistics3D with kernelGetLocalStatistics3D.cl (float
 * mean and unbiased std over the ellipsoid, borders 0), with the block size
 * and radii as the function declares them (GlobalRepetition3D_ passes them
 * in another order). */
int sred_local_stats_3d(sred_ctx *ctx, const float *img, int w, int h, int d, int bw, int bh, int bd, float *means,
                        float *stds);
/* GlobalRepetition3D_.run (GlobalRepetition3D_.java:113-160 ->
 * CLUtils.calculateGlobalRepetitionMap3D :2793-2815) with any of the four
 * metrics on a w x h x d volume (slice z at z*w*h), built as intended where
 * the reference does not run: the statistics above, the 3D relevance mask of
 * Utils.getRelevanceMask3D (noise variance = float mean of std^2 over the
 * inner voxels; 0 where std^2 <= noiseVar * c), nPixels over it, the
 * reference block gathered over the full ellipsoid (the kernels stop one
 * slice short), flat reference blocks scoring 0; per pair the 2D kernels'
 * forms, except NRMSE adding sqrt(mse) * weight (kernelGetGlobalNrmse3D.cl:
 * 111) and AbsDiffStds skipping flat test blocks (kernelGetGlobalAbsDiffStds3D
 * .cl:53); SSIM without the undeclared-name line that stops
 * kernelGetGlobalSsim3D.cl compiling; then S/(Wsum*nPixels + eps), NRMSE and
 * AbsDiffStds inverted where the mask is set, the mask when c > 0 and
 * Utils.normalizeImage3D.  The contraction is the 2D one (precision as
 * sred_global_repetition_2d). */
int sred_global_repetition_3d(sred_ctx *ctx, int metric, int precision, const float *img, int w, int h, int d, int bw,
                              int bh, int bd, float relevance_constant, int normalize, float *out);
int sred_global_repetition_3d_dev(sred_ctx *ctx, int metric, int precision, const float *d_img, int w, int h, int d,
                                  int bw, int bh, int bd, float relevance_constant, int normalize, float *d_out,
                                  float *d_means, float *d_stds, float *d_mask, float *d_S, float *d_Wsum);

/* Utils.getReferenceBlock3D (Utils.java:1047-1155) on a bw x bh x bd block
 * (slice z at z*bw*bh), host only: the ellipsoid's voxels in z, y, x order
 * into px (K floats, K = sred_block_size_3d), normalised
 * (Utils.normalizeArray), mean-subtracted, normalised again; *mean and *std
 * of the result (float sums in order, unbiased std).  Returns K; odd sides
 * of at least 3 (the depth check of :1072-1076). */
int sred_reference_block_3d(const float *block, int bw, int bh, int bd, float *px, float *mean, float *std_);
/* BlockRepetition3D_.run (BlockRepetition3D_.java:86-105 ->
 * CLUtils.calculateBlockRepetitionMap3D :1639-1653) on a normalised
 * w x h x d volume with the reference block of sred_reference_block_3d:
 * the 3D statistics, the kernel of `metric` per inner voxel
 * (kernelGetBlock{Pearson,Ssim,Nrmse,AbsDiffStds}3D.cl; Pearson remaps the
 * test block, subtracts the local mean and remaps again before its dot
 * product), 1/rmse (1 where rmse == 0) for NRMSE and 1 - d for AbsDiffStds,
 * the 3D mask when c > 0 (Utils.getRelevanceMask3D, Utils.applyMask3D) and
 * Utils.normalizeImage3D over it (all inner voxels when c <= 0).  Built as
 * intended where the reference does not run: the block radii, not the
 * block depth (getBlockPearson3D, CLUtils.java:1149) or width
 * (getBlockSsim3D, :1310) that the reference passes.  Errors: block sides
 * larger than the volume's (BlockRepetition3D_.java:97-100). */
int sred_block_repetition_3d(sred_ctx *ctx, int metric, const float *img, int w, int h, int d, const float *tmpl,
                             int bw, int bh, int bd, float tmpl_mean, float tmpl_std, float relevance_constant,
                             int normalize, float *out);
/* Device-pointer form; d_means, d_stds and d_mask (used when c > 0) may be
 * NULL (context scratch) or receive the statistics and the mask. */
int sred_block_repetition_3d_dev(sred_ctx *ctx, int metric, const float *d_img, int w, int h, int d,
                                 const float *tmpl, int bw, int bh, int bd, float tmpl_mean, float tmpl_std,
                                 float relevance_constant, int normalize, float *d_out, float *d_means,
                                 float *d_stds, float *d_mask);

/* The octree GAT estimate of a w x h x d volume: the "Quad/Octree" branch of
 * Utils.getInputImage3D (Utils.java:1234-1254) and the OcTree_ plugin
 * (OcTree_.java:82-96): OcTree_(w, h, d, min_size, alpha).buildTree (a leaf
 * when a side is <= min_size or the F test with (width*height - 1) degrees
 * of freedom says keep; node variance from double sums, pseudo-residuals
 * 7 v - the six neighbours over sqrt(56)), robust means (max_iter), LTS
 * variances (lts_fraction), the leaves in pre-order (children x fastest,
 * then y, then z) and the regression, whose slope g0 and intercept eDC are
 * cast to float (g0_edc[0..1] receive the float values).  boxes (x, y, z,
 * w, h, d), means and variances receive up to cap leaves (each may be
 * NULL); *n_leaves the count.  Errors: a visited node with one pixel per
 * slice (FDistribution(0, 0)), fewer than 3 leaves. */
int sred_octree_gat_3d(sred_ctx *ctx, const float *vol, int w, int h, int d, int min_size, float alpha, int max_iter,
                       float lts_fraction, double *g0_edc, int *n_leaves, int cap, int32_t *boxes, float *means,
                       float *variances);
int sred_octree_gat_3d_dev(sred_ctx *ctx, const float *d_vol, int w, int h, int d, int min_size, float alpha,
                           int max_iter, float lts_fraction, double *g0_edc, int *n_leaves, int cap, int32_t *boxes,
                           float *means, float *variances);
/* OcTree_.applyGATtree (OcTree_.java:471-485) in place on n voxels:
 * ref = 3/8 g0^2 + eDC in double, 0 where v < -ref/g0, else
 * (2/g0) sqrt(g0 v + ref).  Utils.getInputImage3D passes width*height as n
 * (Utils.java:1253), which leaves a one-slice array behind; pass all
 * voxels, as the OcTree_ plugin does. */
int sred_gat_tree_3d(sred_ctx *ctx, float *vol, int64_t n, float g0, float edc);
int sred_gat_tree_3d_dev(sred_ctx *ctx, float *d_vol, int64_t n, float g0, float edc);

/* GATMinimizer3D.userFunction (GATMinimizer3D.java:79-157) at (gain, sigma,
 * offset) on a w x h x d volume: NaN for gain <= 0, sigma < 0 or
 * offset < 0; else the transformed volume (0 where v <= -ref/gain) in
 * blocks of max(w/6,1) x max(h/6,1) x max(d/6,1) voxels, halved along axes
 * with fewer than 2 blocks (:108-127), per block the variance from double
 * sums of v and v*v, and the mean of (var - 1)^2.  The sums run in a fixed
 * order (as sred_gat_objective_2d's).  Errors: a volume for which the
 * reference's block search never ends (a 1-voxel axis while another side
 * stays above 1), more than 4096 blocks. */
int sred_gat_objective_3d(sred_ctx *ctx, const float *vol, int w, int h, int d, double gain, double sigma,
                          double offset, double *error);
/* GATMinimizer3D.run (:40-70): the simplex of sred_gat_minimize_2d (steps
 * 1, 10, 100; NaN ranks worst) over the 3D objective.  The
 * VarianceStabilisingTransform3D_ plugin (VarianceStabilisingTransform3D_
 * .java:70-80) starts it at (gain, sigma, offset) = (1, 100, 10) and applies
 * sred_gat_2d with the fitted doubles to all w*h*d voxels. */
int sred_gat_minimize_3d(sred_ctx *ctx, const float *vol, int w, int h, int d, double gain0, double sigma0,
                         double offset0, int max_iter, double *params, double *error);
int sred_gat_minimize_3d_dev(sred_ctx *ctx, const float *d_vol, int w, int h, int d, double gain0, double sigma0,
                             double offset0, int max_iter, double *params, double *error);

/* ---- relevance-mask tools (RelevanceMask2D_, OptimiseRelevanceMask_,
 *      CalculateLocalMeans_) ---- */
/* The tools' shared input step, in place (RelevanceMask2D_.java:133-164,
 * OptimiseRelevanceMask_.java:124-154, CalculateLocalMeans_.java:163-190):
 * the GAT fit from (gain, sigma, offset) = (0, 100, 0) (as
 * sred_gat_minimize_2d; invalid parameters rank worst), getGAT with the
 * fitted doubles, and (v - min)/(max - min + EPSILON) in double rounded to
 * float.  params[3] (may be NULL) receives the fit.  Synchronises. */
int sred_tool_prep_2d(sred_ctx *ctx, float *img, int w, int h, int max_iter, double *params);
int sred_tool_prep_2d_dev(sred_ctx *ctx, float *d_img, int w, int h, int max_iter, double *params);
/* RelevanceMask2D_ after the input step (RelevanceMask2D_.java:165-250):
 * local standard deviations (kernelGetLocalMeans.cl), the noise variance
 * from double tile sums, threshold = noiseVar * filter_constant
 * (filter_constant alone when noiseVar is 0), mask 0 where std^2 < threshold
 * or std^2 == 0, else 1, 0 on the borders.  Odd square blocks only (the
 * reference swaps the radii and overruns its patch array otherwise).
 * *noise_var and *threshold may be NULL.  Synchronises. */
int sred_relevance_mask_tool_2d(sred_ctx *ctx, const float *img, int w, int h, int bw, int bh, float filter_constant,
                                float *mask, float *noise_var, float *threshold);
int sred_relevance_mask_tool_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh,
                                    float filter_constant, float *d_mask, float *noise_var, float *threshold);
/* OptimiseRelevanceMask_ after the input step (OptimiseRelevanceMask_.java:
 * 155-287): filter constants 0, step, 2 step, ... (a float sum); the
 * plugin's masks compare the squares of kernelGetRelevanceMask2D.cl's local
 * variances with noiseVar * c (c when noiseVar is 0).  *all_ones = the last
 * constant whose mask is all ones inside the borders, *all_zeros = the first
 * whose mask is all zeros.  Errors where the reference loops for ever: step
 * <= 0, a NaN variance, a sum that stops growing, or more than max_steps
 * constants.  Odd square blocks only.  Synchronises. */
int sred_optimise_relevance_mask_2d(sred_ctx *ctx, const float *img, int w, int h, int bw, int bh, float step,
                                    int64_t max_steps, float *all_ones, float *all_zeros);
int sred_optimise_relevance_mask_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh, float step,
                                        int64_t max_steps, float *all_ones, float *all_zeros);
/* RelevanceMaskStack2D_ after the input step (RelevanceMaskStack2D_.java:
 * 140-285): the masks of OptimiseRelevanceMask_.getRelevanceMap (the
 * plugin's local variances squared against noiseVar * c, c when noiseVar is
 * 0) for the constants lower, lower + step, ... (a float sum), (int)
 * ceil((upper - lower) / step) + 1 of them; masks receives them one image
 * after the other, constants (may be NULL) their values, *n_masks the
 * count (also when it exceeds cap, which is an error).  Odd square blocks
 * only.  Synchronises. */
int sred_relevance_mask_stack_2d(sred_ctx *ctx, const float *img, int w, int h, int bw, int bh, float lower,
                                float upper, float step, int cap, float *masks, float *constants, int *n_masks);
int sred_relevance_mask_stack_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh, float lower,
                                    float upper, float step, int cap, float *d_masks, float *constants, int *n_masks);
int sred_simple_regression(const double *x, const double *y, int64_t n, double *slope, double *intercept);

/* ---- all-pairs vector maps (OpenCL kernels without a Java caller) ---- */
/* Per-pixel Hu invariant hu = nu20 + nu02 of kernelGetHuMap.cl:146-172
 * (getInvariant on the bw x bh rectangle, patch = pixel - means[p]), in the
 * kernel's float order; border pixels 0.  means: local means of
 * sred_local_stats_2d (host).  Synchronous. */
int sred_hu_invariants_2d(sred_ctx *ctx, const float *img, const float *means, int w, int h, int bw, int bh,
                          float *out);
/* Global Hu map, kernelGetHuMap.cl:15-172: for every reference pixel i,
 * sum over all interior pixels j of |hu_j - hu_i| * (100^sim - 1) / 99,
 * sim = 1 - |s_i - s_j| / (|s_i| + |s_j| + 1e-7) (getExpDecayWeight,
 * :136-144).  hu as sred_hu_invariants_2d (bit-exact), pair terms in FP32,
 * row sums in FP64.  means/stds: sred_local_stats_2d outputs.  rows/nrows:
 * the reference pixels y*w+x to compute (the kernel's uniqueStdCoords list,
 * :27-39; border and out-of-range entries are skipped); NULL computes every
 * interior pixel.  out (w*h) is 0 at pixels not computed (the kernel adds
 * into a buffer the caller zeroes).  rows is a host array in both forms. */
int sred_global_hu_2d(sred_ctx *ctx, const float *img, const float *means, const float *stds, int w, int h, int bw,
                      int bh, const int32_t *rows, int64_t nrows, float *out);
int sred_global_hu_2d_dev(sred_ctx *ctx, const float *d_img, const float *d_means, const float *d_stds, int w, int h,
                          int bw, int bh, const int32_t *rows, int64_t nrows, float *d_out);
/* Phase correlation map, kernelGetPhaseCorrelationMap.cl:13-179.  That
 * kernel does not compile (nUnique undefined, :27) and its DFT uses integer
 * phases (:66-67); this is textbook phase correlation on the same blocks,
 * FP64: b = (pixel - mean) / (std + 1e-7) on the bw x bh rectangle, unit
 * spectra U = F / (|F| + 1e-7), R = U_i conj(U_j) with i the lower interior
 * index of the pair (one displacement per pair), r = Re IDFT(R); the first
 * strict positive maximum of r in raster order gives the displacement d,
 * wrapped to [-bw/2, bw/2] x [-bh/2, bh/2] (0 if r has no positive value);
 * map[i] = sum over interior j of |d_ij| * w_ij with the kernel's weight
 * (:167-178): 1 if s_i == s_j, else 1 - 2 |s_i - s_j| / (|s_i| + |s_j|).
 * Block sides odd, 3..11.  rows/out as sred_global_hu_2d. */
int sred_phase_correlation_2d(sred_ctx *ctx, const float *img, const float *means, const float *stds, int w, int h,
                              int bw, int bh, const int32_t *rows, int64_t nrows, float *out);
int sred_phase_correlation_2d_dev(sred_ctx *ctx, const float *d_img, const float *d_means, const float *d_stds,
                                  int w, int h, int bw, int bh, const int32_t *rows, int64_t nrows, float *d_out);

/* ---- differentiation (no counterpart in the reference) ---- */
/* Gradient of L = sum_i dL_dmap[i] * map[i] of the global Pearson map
 * (relevance constant 0, normalised output; the forward of
 * sred_global_map_2d_dev with the split-BF16 contraction) with respect to
 * the normalised input pixels, by a hand-written adjoint.  Held fixed: the
 * noise variance, nPixels, the relevance mask and the output min/max; flat
 * blocks (std == 0) pass no gradient through their statistics.  K <= 160
 * (the pair pass in split BF16 on MFMA). */
int sred_global_pearson_adjoint_2d(sred_ctx *ctx, const float *img, int w, int h, int bw, int bh,
                                   const float *dL_dmap, float *dL_dimg);
/* Same on device buffers; d_map (optional) receives the forward map. */
int sred_global_pearson_adjoint_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh,
                                       const float *d_dmap, float *d_grad, float *d_map);

/* ---- introspection for measurement ---- */
/* Launch plan of the global contraction for n_valid valid blocks of
 * block_pixels (K) pixels (no counterpart in the reference; host only).
 * out[8]: padded block count, rows per panel (symmetric) or per workgroup,
 * columns per chunk, chunks, workgroups, partial-sum slots, block pairs
 * evaluated, flags (1 = MFMA, 2 = symmetric, 4 = blocks of K = 32m + 1
 * reflected into K - 1 dimensions). */
int sred_global_plan_2d(int n_valid, int block_pixels, int metric, int precision, int symmetric, int chunk_cols,
                        int64_t *out);
/* Duration (ms, HIP events on the context's stream) and launch count of the
 * dominant kernel of the last call on this context -- the contraction of a
 * global call, the stencil of a block or Hu call, the pair kernel of an
 * adjoint call, the objective's row kernel of a GAT fit (summed over its
 * launches) -- and its work units (block pairs evaluated, or pixels).
 * No counterpart in the reference (measurement). */
int sred_ctx_last_contraction(sred_ctx *ctx, float *ms, int *launches, int64_t *pairs_evaluated);
/* Ordered block pairs whose evaluation the last global call (single or
 * batched, split-BF16 symmetric contraction, Pearson or SSIM) skipped because
 * their t = sigma * sqrt(log2(e) / (h + eps)) differ by more than sqrt(40)
 * (Pearson) or sqrt(47) (SSIM, whose luminance factor reaches 100.5), so that
 * every skipped term is below 2^-40; 0 for other paths.  Synchronises.  No counterpart in the reference (measurement). */
int sred_ctx_last_skipped(sred_ctx *ctx, int64_t *pairs_skipped);

#ifdef __cplusplus
}
#endif
#endif
