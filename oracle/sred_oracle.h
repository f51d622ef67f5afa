/*
 * sred_oracle.h -- CPU restatement of SReD's 2D block-correlation path.
 *
 * TEST INFRASTRUCTURE ONLY.  This library is the checker: the parity tests,
 * the smoke check and the CPU-baseline leg of the benchmark load it to
 * compare against the HIP library (sred_amd/libsred_hip.so).  No product code
 * path links, loads or calls it.
 *
 * Two forms are restated from the reference's text (Java host +
 * OpenCL C kernels):
 *
 *   refsim  -- float32-faithful: the kernels' loop orders, sequential float
 *              accumulation, the double promotions caused by the kernels'
 *              unsuffixed macro literals, and the Java float host steps.
 *              This is "the reference's own map".
 *   oracle64 -- the same mathematics in FP64 with compensated sums; discrete
 *              decisions (relevance mask, threshold skip, sigma==0 special
 *              cases) are taken from the float32 statistics exactly as the
 *              reference takes them, magnitudes are FP64.
 *
 * Every function cites the reference file:line it follows (paths relative to
 * the SReD repository root).  Pinning of each function is described in the
 * test files under tests/ (known-answer vectors from the reference's own
 * JUnit tests, closed forms, invariants, brute force on tiny inputs).
 */
#ifndef SRED_ORACLE_H
#define SRED_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { OR_PEARSON = 0, OR_SSIM = 1, OR_NRMSE = 2, OR_ABSDIFF = 3 };

/* Largest footprint the block-vector restatements hold per thread (the HIP
 * library's cap: a 63 x 63 ellipse has 3109 pixels).  Larger footprints are
 * refused: NaN outputs from the void entry points, -1 from the others. */
#define OR_MAX_K 3200

/* Elliptical footprint, GlobalRepetition2D_.java:102-111 and the kernel
 * predicate kernelGetGlobalPearson2D.cl:44-50.  Writes the row-major offsets
 * (dx, dy) of the K member pixels if dx/dy are non-NULL; returns K. */
int or_footprint(int bw, int bh, int *dx, int *dy);

/* The double an OpenCL compiler reads from Java's ""+float rendering of f
 * (CLUtils.java:1705-1712): the shortest decimal that round-trips to f,
 * parsed as double. */
double or_java_float_literal(float f);

/* Utils.getInputImage2D min-max normalisation, Utils.java:708-722. */
void or_normalize_input(float *img, int64_t n);

/* Utils.getReferenceBlock2D, Utils.java:557-639 (including the double
 * mean subtraction in the std, Utils.java:626-636).  px receives the K
 * mean-subtracted template pixels.  Returns K, or -1 for even sizes. */
int or_reference_block(const float *blk, int bw, int bh, float *px, float *mean, float *std_);

/* kernelGetLocalStatistics2D.cl:9-55 in float32 (sequential sums, border 0). */
void or_local_stats_f32(const float *img, int w, int h, int bw, int bh, float *means, float *stds);
/* The same statistics in FP64 (mean and unbiased variance). */
void or_local_stats_f64(const float *img, int w, int h, int bw, int bh, double *means, double *vars);

/* Utils.getMeanAndVarBlock2D, Utils.java:854-900 (one-pass float).  Returns
 * 0, or -1 for invalid coordinates / non-finite input (the reference throws). */
int or_mean_var_block(const float *img, int w, int h, int x0, int y0, int x1, int y1, float *mean, float *var);
/* Utils.getMeanNoiseVar2D, Utils.java:790-837. */
float or_noise_var(const float *img, int w, int h);
/* Utils.getRelevanceMask, Utils.java:944-968, plus the nPixels count of
 * GlobalRepetition2D_.java:133-140.  mask is W*H {0,1}. */
void or_relevance_mask(const float *stds, int w, int h, int brw, int brh, float noise_var, float c,
                       float *mask, float *threshold, float *n_pixels);
/* Utils.applyMask2D, Utils.java:915-925. */
void or_apply_mask(float *map, int64_t n, const float *mask);
/* Utils.normalizeImage2D, Utils.java:985-1031, in place (mask may be NULL).
 * Returns 0, or -1 for a non-binary mask (the reference throws). */
int or_normalize_map(float *map, int w, int h, int brw, int brh, const float *mask);
int or_normalize_map_f64(double *map, int w, int h, int brw, int brh, const float *mask);

/* Global accumulators S (similarity sum) and Wsum (weight sum), float32,
 * kernelGetGlobal{Pearson,Ssim,Nrmse,AbsDiffStds}2D.cl.  filter_param and
 * threshold are the float values the host renders into the kernel macros.
 * rows: optional list of reference pixel indices (y*w+x) to compute; NULL
 * computes every pixel.  S/Wsum are W*H (only listed pixels are written).
 * nthreads <= 0 uses all OpenMP threads. */
void or_global_accum_f32(int metric, const float *img, const float *means, const float *stds, int w, int h,
                         int bw, int bh, float filter_param, float threshold, float *S, float *Wsum,
                         const int32_t *rows, int64_t nrows, int nthreads);
/* FP64 accumulators; stats recomputed in FP64; validity and sigma==0
 * decisions from the float32 stds (stds32). */
void or_global_accum_f64(int metric, const float *img, const float *stds32, int w, int h, int bw, int bh,
                         float filter_param, float threshold, double *S, double *Wsum,
                         const int32_t *rows, int64_t nrows, int nthreads);

/* Host finalize of CLUtils.getGlobal*2D (CLUtils.java:1747-1770, 1983-2020,
 * 2112-2149): rep = S/(Wsum*nPixels+eps) on the interior, inversion for
 * NRMSE/AbsDiff where mask>0, mask if c>0, masked normalisation. */
int or_global_finalize(int metric, const float *S, const float *Wsum, const float *mask, float n_pixels,
                       float c, int normalize, int w, int h, int bw, int bh, float *out);
int or_global_finalize_f64(int metric, const double *S, const double *Wsum, const float *mask, double n_pixels,
                           float c, int normalize, int w, int h, int bw, int bh, double *out);

/* Full global-repetition pipeline from a normalised input image:
 * GlobalRepetition2D_.java:124-145 with VST off.  fp64 selects oracle64. */
int or_global_map(int metric, int fp64, const float *img, int w, int h, int bw, int bh, float c, int normalize,
                  float *out32, double *out64, int nthreads);

/* Block repetition, CLUtils.getBlock{Pearson,Ssim,Nrmse,AbsDiffStds}2D
 * (CLUtils.java:486-981) with their kernels.  tmpl is the K template pixels
 * as handed to the kernel (ReferenceBlock2D.getPixels()). */
int or_block_map(int metric, const float *img, int w, int h, const float *tmpl, int bw, int bh, float tmpl_mean,
                 float tmpl_std, float c, int normalize, float *out);
/* Raw per-window kernel output only (before any host post-processing). */
int or_block_raw(int metric, const float *img, const float *means, const float *stds, int w, int h,
                 const float *tmpl, int bw, int bh, float tmpl_mean, float tmpl_std, float *out);

/* Hu template metric, kernelGetPatchHu2D.cl:11-282 (no Java caller in the
 * reference).  FP64 throughout, following the kernel's loops:
 *   - template: the bw x bh rectangle `tmpl` is used as given where the float
 *     ellipse predicate holds (:36-44); its moments use the double predicate
 *     (:56-67, :80-96);
 *   - window: mean over the square minus its four corners (integer division
 *     in the predicate, :129-140) divided by K; mean subtraction, min/max and
 *     min-max normalisation over the float ellipse (:143-182);
 *   - raw moments with 1-based coordinates (i+1, j+1), central moments with
 *     0-based ones (:190-234); hu1 = nu20 + nu02, nu = mu / (M00^2 + eps);
 *   - out = (float) sqrt(hu1_template * hu1_window) (:267); border pixels 0.
 * pow(x, 2.0) is x*x (the correctly rounded square).  eps is the double the
 * kernel's EPSILON macro denotes.  Only hu1 reaches the output. */
double or_hu1_template(const float *tmpl, int bw, int bh, double eps);
void or_block_hu(const float *img, int w, int h, const float *tmpl, int bw, int bh, double eps, float *out);
/* The template preparation this build defines for the Hu map (the kernel
 * expects "mean-subtracted and normalised" template pixels and has no
 * caller): the window steps of :129-182 applied to the template rectangle
 * (mean over the square minus corners / K, subtraction and min-max over the
 * float ellipse), in FP64, rounded to float; pixels outside the ellipse 0. */
void or_hu_prepare_template(const float *blk, int bw, int bh, double eps, float *out);

/* Global Hu map, kernelGetHuMap.cl:15-172 (no caller in the reference).
 * or_hu_invariant: hu = inv20 + inv02 of getInvariant (:146-172) on the
 * bw x bh rectangle around (x0, y0) with patch = pixel - mean, in the
 * kernel's float order (see the definition for the promotions).
 * or_hu_invariants: that value at every interior pixel (border 0).
 * or_hu_weight: getExpDecayWeight (:136-144) in FP64.
 * or_global_hu_sums: map[i] = sum over interior j of |hu_j - hu_i| w(s_i, s_j)
 * (:87-133) in FP64 for the interior pixels of rows (NULL: all), others 0. */
float or_hu_invariant(const float *img, int w, int bw, int bh, int x0, int y0, float mean, double eps);
void or_hu_invariants(const float *img, const float *means, int w, int h, int bw, int bh, double eps, float *hu);
double or_hu_weight(float a, float b, double eps);
void or_global_hu_sums(const float *hu, const float *stds, int w, int h, int bw, int bh, double eps,
                       const int32_t *rows, int64_t nrows, double *out, int nthreads);

/* Phase correlation map, kernelGetPhaseCorrelationMap.cl:13-179.  The
 * reference kernel does not compile; this is the textbook specification on
 * the same blocks and weight (full statement in sred_oracle.c):
 * normalised blocks (pixel - mean) / (std + eps) on the bw x bh rectangle,
 * 2D DFTs F, unit spectra U = F / (|F| + eps), R = U_i conj(U_j),
 * r = Re IDFT(R), the
 * first positive strict maximum in raster order, its coordinates wrapped to
 * signed displacements, map[i] = sum_j |d_ij| w(s_i, s_j) in FP64.
 * or_phase_peak_blocks: r (n values, may be NULL) and the wrapped peak of two
 * normalised blocks; returns 1, or 0 if no value of r is positive.
 * or_phase_weight: getExpDecayWeight (:167-178), float, as double. */
int or_phase_peak_blocks(const double *bi, const double *bj, int bw, int bh, double eps, double *r_out, int *dx,
                         int *dy);
double or_phase_weight(float a, float b);
void or_phase_map(const float *img, const float *means, const float *stds, int w, int h, int bw, int bh,
                  double eps, const int32_t *rows, int64_t nrows, double *out, int nthreads);

/* Variance stabilisation (VarianceStabilisingTransform2D_.getGAT and
 * GATMinimizer2D).  or_gat: getGAT in place (:136-151).  or_gat_objective:
 * GATMinimizer2D.userFunction with a fixed 64-lane summation order
 * (definition in sred_oracle.c).  or_nelder_mead3 / or_gat_minimize: the build's own
 * Nelder-Mead simplex in place of ImageJ's Minimizer (parity of the fit
 * with ImageJ unpinned; the library's simplex follows this one step for
 * step). */
void or_gat(float *px, int64_t n, double gain, double sigma, double offset);
double or_gat_objective(const float *px, int w, int h, double gain, double sigma, double offset);
typedef double (*or_fn3)(const double *x, void *ctx);
double or_nelder_mead3(or_fn3 f, void *ctx, const double *x0, const double *d, int max_iter, double *x);
double or_gat_minimize(const float *px, int w, int h, double gain0, double sigma0, double offset0, int max_iter,
                       double *params);

/* Quadtree GAT estimate (QuadTree_.java:135-427 as called by
 * Utils.getInputImage2D, Utils.java:686-703).  Float arithmetic and sum
 * orders as the reference; definitions and statuses in sred_oracle.c.
 * or_beta_reg / or_f_quantile / or_simple_regression restate commons-math3
 * 3.6.1 (the reference's dependency, not in the reference tree). */
double or_beta_reg(double a, double b, double x);
double or_f_quantile(double d1, double d2, double p);
int or_simple_regression(const double *x, const double *y, int64_t n, double *slope, double *intercept);
void or_qt_criterion(const float *img, int w, int h, int x0, int y0, int nw, int nh, float *block_var,
                     float *noise_var);
float or_qt_robust_mean(const float *img, int w, int x0, int y0, int nw, int nh, int max_iter);
float or_qt_lts_variance(const float *img, int w, int x0, int y0, int nw, int nh, float mean, float frac);
int or_quadtree_gat(const float *img, int w, int h, int min_size, float alpha, int max_iter, float frac, int cap,
                    int32_t *rects, float *means, float *vars, int *nleaves, double *g0, double *edc);
void or_gat_tree(float *px, int64_t n, double g0, double edc);

/* Volumes: GlobalRepetition3D_ with Pearson's R as intended (definitions
 * and the decisions on the reference's own defects in sred_oracle.c). */
int or_footprint3d(int bw, int bh, int bd, int *dx, int *dy, int *dz);
void or_local_stats3d_f32(const float *img, int w, int h, int d, int bw, int bh, int bd, float *means, float *stds);
float or_vol_noise_var(const float *stds, int w, int h, int d, int brw, int brh, int brd);
int or_global_pearson3d(const float *img, int w, int h, int d, int bw, int bh, int bd, float c, int normalize,
                        int fp64, double *out, float *S_out, float *W_out, float *mask_out, float *noise_var,
                        float *n_pixels);
int or_global_repetition3d(int metric, const float *img, int w, int h, int d, int bw, int bh, int bd, float c,
                           int normalize, int fp64, double *out, float *S_out, float *W_out, float *mask_out,
                           float *noise_var, float *n_pixels);
/* BlockRepetition3D_ (Utils.getReferenceBlock3D, the four block kernels and
 * their host steps), as intended; definitions in sred_oracle.c. */
int or_reference_block3d(const float *blk, int bw, int bh, int bd, float *px, float *mean, float *std_);
int or_block_raw3d(int metric, const float *img, const float *means, const float *stds, int w, int h, int d,
                   const float *tmpl, int bw, int bh, int bd, float tmpl_mean, float tmpl_std, float *out);
int or_block_map3d(int metric, const float *img, int w, int h, int d, const float *tmpl, int bw, int bh, int bd,
                   float tmpl_mean, float tmpl_std, float c, int normalize, float *out);
/* OcTree_ GAT estimate of a volume and its transform; definitions in
 * sred_oracle.c. */
void or_oc_criterion(const float *img, int w, int h, int d, int x0, int y0, int z0, int nw, int nh, int nd,
                     float *block_var, float *noise_var);
float or_oc_robust_mean(const float *img, int w, int h, int x0, int y0, int z0, int nw, int nh, int nd,
                        int max_iter);
float or_oc_lts_variance(const float *img, int w, int h, int x0, int y0, int z0, int nw, int nh, int nd, float mean,
                         float frac);
int or_octree_gat(const float *img, int w, int h, int d, int min_size, float alpha, int max_iter, float frac,
                  int cap, int32_t *boxes, float *means, float *vars, int *nleaves, double *g0, double *edc);
void or_gat_tree3d(float *px, int64_t n, float g0, float edc);
/* GATMinimizer3D: block grid, objective (device sum order) and the simplex. */
int or_gat_blocks3d(int w, int h, int d, int *bw, int *bh, int *bd, int *nbx, int *nby, int *nbz);
double or_gat_objective3d(const float *px, int w, int h, int d, double gain, double sigma, double offset);
double or_gat_minimize3d(const float *px, int w, int h, int d, double gain0, double sigma0, double offset0,
                         int max_iter, double *params);

/* Relevance-mask tools (RelevanceMask2D_, OptimiseRelevanceMask_, the
 * intent of CalculateLocalMeans_); definitions in sred_oracle.c. */
void or_tool_prep(float *img, int w, int h, int max_iter, double *params);
float or_noise_var_double(const float *img, int w, int h);
int or_relevance_mask_tool(const float *img, int w, int h, int bw, int bh, float c, float *mask, float *noise_var,
                           float *threshold);
void or_local_var_f32(const float *img, int w, int h, int bw, int bh, float *var);
int or_optimise_relevance_mask(const float *img, int w, int h, int bw, int bh, float step, int max_steps,
                               float *all_ones, float *all_zeros);
int or_relevance_mask_stack(const float *img, int w, int h, int bw, int bh, float lower, float upper, float step,
                            int cap, float *masks, float *constants, int *n);

#ifdef __cplusplus
}
#endif
#endif
