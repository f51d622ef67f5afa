// block.hip -- block repetition: one template against every window.
//
// Reference: kernelGetBlock{Pearson,Ssim,Nrmse,AbsDiffStds}2D.cl, launched
// over the whole image by CLUtils.getBlock*2D (CLUtils.java:486-981), after
// kernelGetLocalStatistics2D.  Here statistics and comparison are one
// LDS-tiled stencil: each workgroup stages its input tile with halo once,
// every pixel computes its window mean/std in the reference's float order
// (identical to the statistics kernel) and the template comparison from the
// same LDS tile; the template sits in LDS too.  Window statistics are written
// only when the caller needs them (relevance mask).
#include <float.h>

#include <algorithm>
#include <cmath>

#include "sred_internal.h"

namespace sred {

constexpr int kBlkTile = 16;

__global__ __launch_bounds__(256) void k_block(int metric, const float *__restrict__ img, int w, int h, int brw,
                                               int brh, const int *__restrict__ offs, int K,
                                               const float *__restrict__ tmpl, double tmpl_mean_d, double tmpl_std_d,
                                               float tmpl_std_f, double eps_d, float *__restrict__ out,
                                               float *__restrict__ means_out, float *__restrict__ stds_out)
{
#pragma clang fp contract(off)
    extern __shared__ float sm[];
    const int tw = kBlkTile + 2 * brw, th = kBlkTile + 2 * brh;
    float *tile = sm;
    float *tp = sm + tw * th;
    const int x0 = blockIdx.x * kBlkTile, y0 = blockIdx.y * kBlkTile;
    for (int i = threadIdx.x; i < tw * th; i += blockDim.x) {
        int ly = i / tw, lx = i - ly * tw;
        int gy = y0 - brh + ly, gx = x0 - brw + lx;
        tile[i] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? img[(int64_t)gy * w + gx] : 0.0f;
    }
    for (int i = threadIdx.x; i < K; i += blockDim.x) tp[i] = tmpl[i];
    __syncthreads();
    const int tx = threadIdx.x % kBlkTile, ty = threadIdx.x / kBlkTile;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx >= w || gy >= h) return;
    const int64_t o = (int64_t)gy * w + gx;
    if (gx < brw || gx >= w - brw || gy < brh || gy >= h - brh) {
        out[o] = 0.0f;
        if (means_out) { means_out[o] = 0.0f; stds_out[o] = 0.0f; }
        return;
    }
    const float *c = tile + (ty + brh) * tw + (tx + brw);
    // window statistics, kernelGetLocalStatistics2D.cl:40-55
    float m = 0.0f;
    for (int k = 0; k < K; k++) m += c[offs[k]];
    m /= (float)K;
    float v = 0.0f;
    for (int k = 0; k < K; k++) {
        float d = c[offs[k]] - m;
        v += d * d;
    }
    v /= (float)(K - 1);
    const float cs = sqrtf(v);
    if (means_out) { means_out[o] = m; stds_out[o] = cs; }

    float r;
    if (metric == SRED_ABSDIFF) {  // kernelGetBlockAbsDiffStds2D.cl:103-105
        r = fabsf(tmpl_std_f - cs);
    } else if (metric == SRED_NRMSE) {  // kernelGetBlockNrmse2D.cl:69-74
        float mse = 0.0f;
        for (int k = 0; k < K; k++) {
            float e = tp[k] - (c[offs[k]] - m);
            mse += e * e;
        }
        mse /= (float)K;
        r = sqrtf(mse);
    } else {
        float cov = 0.0f;  // kernelGetBlockPearson2D.cl:70-74
        for (int k = 0; k < K; k++) cov += tp[k] * (c[offs[k]] - m);
        cov /= (float)(K - 1);
        if (tmpl_std_d == 0.0 && cs == 0.0f) {
            r = 1.0f;
        } else if (tmpl_std_d == 0.0 || cs == 0.0f) {
            r = 0.0f;
        } else if (metric == SRED_PEARSON) {  // kernelGetBlockPearson2D.cl:88 (double literals)
            r = fmaxf(0.0f, (float)((double)cov / ((tmpl_std_d * (double)cs) + eps_d)));
        } else {  // kernelGetBlockSsim2D.cl:169 (sum denominator, double literals)
            const float c1 = 0.0001f, c2 = 0.0009f;
            double num = (2.0 * tmpl_mean_d * (double)m + (double)c1) * (double)(2.0f * cov + c2);
            double den = ((tmpl_mean_d * tmpl_mean_d) + (double)(m * m) + (double)c1) *
                         ((tmpl_std_d * tmpl_std_d) + (double)(cs * cs) + (double)c2);
            r = fmaxf(0.0f, (float)(num / den));
        }
    }
    out[o] = r;
}

int block_raw_dev(sred_ctx *ctx, const Geometry &g, int metric, const float *d_img, float *d_means_out,
                  float *d_stds_out, const float *h_tmpl, float tmpl_mean, float tmpl_std, float *d_out)
{
    if (metric < 0 || metric > 3) {
        set_error("unknown metric");
        return SRED_ERR_UNKNOWN_METRIC;
    }
    const int tw = kBlkTile + 2 * g.brw, th = kBlkTile + 2 * g.brh;
    SRED_TRY(upload_offsets(ctx, g, tw, &ctx->offs, ctx->offs_key));
    SRED_TRY(ctx->tmpl.ensure(sizeof(float) * g.K));
    SRED_HIP_CHECK(hipMemcpy(ctx->tmpl.ptr, h_tmpl, sizeof(float) * g.K, hipMemcpyHostToDevice));
    dim3 grid((g.w + kBlkTile - 1) / kBlkTile, (g.h + kBlkTile - 1) / kBlkTile);
    size_t lds = sizeof(float) * ((size_t)tw * th + g.K);
    SRED_TRY(time_begin(ctx));
    hipLaunchKernelGGL(k_block, grid, dim3(256), lds, ctx->stream, metric, d_img, g.w, g.h, g.brw, g.brh,
                       ctx->offs.as<int>(), g.K, ctx->tmpl.as<float>(), java_float_literal(tmpl_mean),
                       java_float_literal(tmpl_std), tmpl_std, java_float_literal(kEps), d_out,
                       d_stds_out ? d_means_out : nullptr, d_means_out ? d_stds_out : nullptr);
    SRED_HIP_CHECK(hipGetLastError());
    SRED_TRY(time_end(ctx, g.n()));
    return SRED_OK;
}

}  // namespace sred
