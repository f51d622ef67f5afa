// tools.hip -- the relevance-mask tools: RelevanceMask2D_ (the mask alone),
// OptimiseRelevanceMask_ (the range of useful filter constants) and the
// intent of CalculateLocalMeans_ (local statistics), with their shared input
// step.
//
// * tool_prep_dev: the three plugins' input step (RelevanceMask2D_.java:
//   133-164, OptimiseRelevanceMask_.java:124-154, CalculateLocalMeans_.java:
//   163-190): GATMinimizer2D from (0, 100, 0) (gain 0 is invalid: the
//   simplex ranks NaN worst), getGAT with the fitted doubles (not rounded to
//   float, unlike getInputImage2D), then (v - min) / (max - min + EPSILON) in
//   double, rounded to float (k_tool_remap; min and max exact).
// * relevance_mask_tool_dev: RelevanceMask2D_ after the input step: local
//   standard deviations (kernelGetLocalMeans.cl computes exactly what
//   kernelGetLocalStatistics2D.cl does), the noise variance from double tile
//   sums, threshold = noiseVar * c (c when noiseVar is 0), and 0 where
//   std^2 < threshold or std^2 == 0 (k_tool_mask).
// * optimise_relevance_mask_dev: OptimiseRelevanceMask_'s scan of c = 0,
//   step, 2 step, ... (a float sum) for the last all-ones and the first
//   all-zeros mask.  The plugin's "standard deviations" come from
//   kernelGetRelevanceMask2D.cl, which writes the variance, so the masks
//   compare variance^2.  A mask is all ones iff no variance^2 is 0 and the
//   smallest is >= threshold, all zeros iff every nonzero variance^2 is
//   below the threshold: one reduction (k_var2_extremes) gives both for
//   every c, and the scan runs on the host without rebuilding masks.
// Blocks must be odd and square: the plugins swap the radii (bRW = bH/2)
// and size their patch arrays with the swapped loops, which overruns them
// for other shapes.
#include <cfloat>
#include <cmath>
#include <cstring>

#include "sred_internal.h"

namespace sred {

namespace {

__global__ __launch_bounds__(256) void k_tool_remap(float *__restrict__ x, int64_t n, const unsigned int *__restrict__ mm)
{
#pragma clang fp contract(off)
    const double mn = (double)ord2f(mm[0]), mx = (double)ord2f(mm[1]);
    const double den = mx - mn + (double)kEps;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        x[i] = (float)(((double)x[i] - mn) / den);
}

__global__ __launch_bounds__(256) void k_tool_mask(const float *__restrict__ stds, int w, int h, int br, float thr,
                                                   float *__restrict__ mask)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const int64_t o = (int64_t)y * w + x;
    float m = 0.0f;
    if (x >= br && x < w - br && y >= br && y < h - br) {
        const float v = stds[o] * stds[o];
        m = (v < thr || v == 0.0f) ? 0.0f : 1.0f;
    }
    mask[o] = m;
}

// ext = (min key of the non-NaN variance^2, max key of the nonzero non-NaN
// ones, any zero, any NaN) over the pixels inside the borders
__global__ __launch_bounds__(256) void k_var2_extremes(const float *__restrict__ var, int w, int h, int br,
                                                       unsigned int *__restrict__ ext)
{
#pragma clang fp contract(off)
    __shared__ unsigned int red[4][4];
    unsigned int lo = 0xffffffffu, hi = 0u, zero = 0u, nan = 0u;
    for (int y = br + blockIdx.x; y < h - br; y += gridDim.x) {  // rows inside the borders
        const float *row = var + (int64_t)y * w;
        for (int x = br + threadIdx.x; x < w - br; x += blockDim.x) {
            const float s = row[x];
            const float v = s * s;
            if (v != v) {
                nan = 1u;
                continue;
            }
            const unsigned int k = f2ord(v);
            lo = min(lo, k);
            if (v == 0.0f)
                zero = 1u;
            else
                hi = max(hi, k);
        }
    }
    for (int off = 32; off > 0; off >>= 1) {
        lo = min(lo, (unsigned int)__shfl_xor((int)lo, off));
        hi = max(hi, (unsigned int)__shfl_xor((int)hi, off));
        zero |= (unsigned int)__shfl_xor((int)zero, off);
        nan |= (unsigned int)__shfl_xor((int)nan, off);
    }
    const int wv = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) {
        red[wv][0] = lo;
        red[wv][1] = hi;
        red[wv][2] = zero;
        red[wv][3] = nan;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int k = 1; k < 4; k++) {
            lo = min(lo, red[k][0]);
            hi = max(hi, red[k][1]);
            zero |= red[k][2];
            nan |= red[k][3];
        }
        atomicMin(ext, lo);
        atomicMax(ext + 1, hi);
        if (zero) atomicOr(ext + 2, 1u);
        if (nan) atomicOr(ext + 3, 1u);
    }
}

__global__ void k_ext_init(unsigned int *ext)
{
    if (threadIdx.x == 0) {
        ext[0] = 0xffffffffu;
        ext[1] = 0u;
        ext[2] = 0u;
        ext[3] = 0u;
    }
}

float ord2f_host(unsigned int u)
{
    const unsigned int b = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
    float f;
    memcpy(&f, &b, sizeof f);
    return f;
}

int check_square(int w, int h, int bw, int bh, Geometry *g)
{
    if (bw != bh) {
        set_error("the relevance-mask tools need square blocks (the reference swaps the radii)");
        return SRED_ERR_ARG;
    }
    return make_geometry(w, h, bw, bh, g);
}

}  // namespace

int tool_prep_dev(sred_ctx *ctx, float *d_img, int w, int h, int max_iter, double *params)
{
    double p[3];
    SRED_TRY(gat_minimize_dev(ctx, d_img, w, h, 0.0, 100.0, 0.0, max_iter, p, nullptr));
    const int64_t n = (int64_t)w * h;
    SRED_TRY(gat_dev(ctx, d_img, n, p[0], p[1], p[2]));
    SRED_TRY(ctx->seg_mm.ensure(256));
    unsigned int *mm = ctx->seg_mm.as<unsigned int>();
    SRED_TRY(seg_minmax_dev(ctx, d_img, n, 1, mm));
    hipLaunchKernelGGL(k_tool_remap, dim3((int)std::min<int64_t>((n + 255) / 256, 4096)), dim3(256), 0, ctx->stream,
                       d_img, n, mm);
    SRED_HIP_CHECK(hipGetLastError());
    if (params)
        for (int k = 0; k < 3; k++) params[k] = p[k];
    return SRED_OK;
}

int relevance_mask_tool_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh, float c, float *d_mask,
                            float *noise_var, float *threshold)
{
#pragma clang fp contract(off)
    Geometry g;
    SRED_TRY(check_square(w, h, bw, bh, &g));
    SRED_TRY(ctx->means.ensure(sizeof(float) * g.n()));
    SRED_TRY(ctx->stds.ensure(sizeof(float) * g.n()));
    SRED_TRY(launch_local_stats(ctx, g, d_img, ctx->means.as<float>(), ctx->stds.as<float>()));
    float nv = 0.0f;
    SRED_TRY(noise_var_dev(ctx, d_img, w, h, &nv, true));
    const float thr = (nv == 0.0f) ? c : nv * c;  // RelevanceMask2D_.java:236-244
    hipLaunchKernelGGL(k_tool_mask, dim3((w + 255) / 256, h), dim3(256), 0, ctx->stream, ctx->stds.as<float>(), w, h,
                       g.brw, thr, d_mask);
    SRED_HIP_CHECK(hipGetLastError());
    if (noise_var) *noise_var = nv;
    if (threshold) *threshold = thr;
    return SRED_OK;
}

int optimise_relevance_mask_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh, float step,
                                int64_t max_steps, float *all_ones, float *all_zeros)
{
#pragma clang fp contract(off)
    Geometry g;
    SRED_TRY(check_square(w, h, bw, bh, &g));
    if (!(step > 0.0f) || max_steps < 1) {
        set_error("the filter-constant step must be positive (the reference loops for ever otherwise)");
        return SRED_ERR_ARG;
    }
    SRED_TRY(ctx->means.ensure(sizeof(float) * g.n()));
    SRED_TRY(ctx->stds.ensure(sizeof(float) * g.n()));
    SRED_TRY(launch_local_stats(ctx, g, d_img, ctx->means.as<float>(), ctx->stds.as<float>(), true));
    SRED_TRY(ctx->seg_mm.ensure(256));
    unsigned int *ext = ctx->seg_mm.as<unsigned int>();
    hipLaunchKernelGGL(k_ext_init, dim3(1), dim3(64), 0, ctx->stream, ext);
    const int rows = h - 2 * g.brh;
    if (rows > 0 && w - 2 * g.brw > 0)
        hipLaunchKernelGGL(k_var2_extremes, dim3(std::min(rows, 512)), dim3(256), 0, ctx->stream,
                           ctx->stds.as<float>(), w, h, g.brw, ext);
    SRED_HIP_CHECK(hipGetLastError());
    float nv = 0.0f;
    SRED_TRY(noise_var_dev(ctx, d_img, w, h, &nv));  // OptimiseRelevanceMask_.getRelevanceMap's float tile sums
    unsigned int e[4];
    SRED_HIP_CHECK(hipMemcpyAsync(e, ext, sizeof e, hipMemcpyDeviceToHost, ctx->stream));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (e[3]) {
        set_error("a local variance is NaN: no constant gives an all-zero mask (the reference loops for ever)");
        return SRED_ERR_NONFINITE;
    }
    const bool any = e[0] != 0xffffffffu, any_nonzero = e[1] != 0u, any_zero = e[2] != 0u;
    const float vmin = any ? ord2f_host(e[0]) : 0.0f, vmax = any_nonzero ? ord2f_host(e[1]) : 0.0f;
    float cst = 0.0f, ones = 0.0f;
    for (int64_t it = 0; it < max_steps; it++) {
        const float thr = (nv == 0.0f) ? cst : nv * cst;
        const bool every_one = !any_zero && (!any || vmin >= thr);
        const bool every_zero = !any_nonzero || vmax < thr;
        if (every_one) ones = cst;
        if (every_zero) {
            *all_ones = ones;
            *all_zeros = cst;
            return SRED_OK;
        }
        const float next = cst + step;
        if (next == cst) break;  // the float sum no longer moves: the reference loops for ever
        cst = next;
    }
    set_error("no all-zero mask within the step limit");
    return SRED_ERR_ARG;
}

int relevance_mask_stack_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh, float lower,
                             float upper, float step, int cap, float *d_masks, float *constants, int *n_masks)
{
#pragma clang fp contract(off)
    // RelevanceMaskStack2D_ after the input step (RelevanceMaskStack2D_.java:
    // 140-285): the plugin's local variances (kernelGetRelevanceMask2D.cl),
    // then one OptimiseRelevanceMask_.getRelevanceMap mask per constant
    // lower, lower + step, ... (a float sum), (int) ceil((upper - lower) /
    // step) + 1 of them (:262-263)
    *n_masks = 0;
    Geometry g;
    SRED_TRY(check_square(w, h, bw, bh, &g));
    const double q = std::ceil((double)((upper - lower) / step));
    if (!(step > 0.0f) || !(q >= 0.0) || q > 1e8) {
        set_error("the filter constants need a positive step and upper >= lower");
        return SRED_ERR_ARG;
    }
    const int count = (int)q + 1;
    *n_masks = count;
    if (count > cap) {
        set_error("more masks than the output holds");
        return SRED_ERR_ARG;
    }
    SRED_TRY(ctx->means.ensure(sizeof(float) * g.n()));
    SRED_TRY(ctx->stds.ensure(sizeof(float) * g.n()));
    SRED_TRY(launch_local_stats(ctx, g, d_img, ctx->means.as<float>(), ctx->stds.as<float>(), true));
    float nv = 0.0f;
    SRED_TRY(noise_var_dev(ctx, d_img, w, h, &nv));  // getRelevanceMap's float tile sums
    float c = lower;
    for (int i = 0; i < count; i++) {
        const float thr = (nv == 0.0f) ? c : nv * c;
        hipLaunchKernelGGL(k_tool_mask, dim3((w + 255) / 256, h), dim3(256), 0, ctx->stream, ctx->stds.as<float>(), w,
                           h, g.brw, thr, d_masks + (size_t)i * g.n());
        if (constants) constants[i] = c;
        c += step;
    }
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

}  // namespace sred
