// block3d.hip -- block repetition of volumes: BlockRepetition3D_
// (BlockRepetition3D_.java:32-115) -> CLUtils.calculateBlockRepetitionMap3D
// (:1639-1653) -> getBlock{Pearson,Ssim,Nrmse,AbsDiffStds}3D (:1133-1637)
// with kernelGetBlock*3D.cl, as intended where the reference does not run:
//   * getBlockPearson3D passes the block depth as the z radius (:1149), so
//     its statistics and kernel treat every voxel as border (an all-zero
//     map): the radius is used;
//   * getBlockSsim3D passes the block width as the x radius (:1310), so its
//     kernel gathers more voxels than its block array holds: the radius is
//     used.
// One thread per voxel, 64 x 4 voxels per workgroup (x across the wave, so
// each footprint offset is one coalesced row read per wave); the offsets and
// the template are indexed uniformly (scalar loads).  Pearson reads its K
// voxels in three passes (min/max; remap, local-mean subtraction, min/max;
// the second remap and the dot product): up to kCacheK voxels the second
// pass's values wait in LDS for the third, beyond it they are recomputed
// bit for bit.  The macro literals of
// the kernels are double (ref_mean, ref_std, EPSILON), so the remaps divide
// in double, Pearson's quotient is double and SSIM's numerator is rounded
// to float before its double division (kernelGetBlockSsim3D.cl:101).  The
// host steps that follow the kernels (1/rmse, 1 where rmse == 0, for NRMSE:
// CLUtils.java:1459-1472; 1 - d for AbsDiffStds: :1597-1605) are fused.
#include <algorithm>
#include <cfloat>
#include <cmath>
#include <vector>

#include "sred_internal.h"

namespace sred {

namespace {

constexpr int kCacheK = 48;  // Pearson keeps the remapped block in LDS up to this size (48 KB per workgroup)

// CACHE (Pearson, K <= kCacheK): pass 2 parks each lane's remapped,
// mean-subtracted values in LDS ([k][thread], no bank conflicts) and pass 3
// reads them back instead of gathering and dividing again.
template <int M, bool CACHE>
__global__ __launch_bounds__(256) void k_block3d(const float *__restrict__ img, const float *__restrict__ means,
                                                 const float *__restrict__ stds, int w, int h, int d, int brw,
                                                 int brh, int brd, const int *__restrict__ offs,
                                                 const float *__restrict__ tmpl, int K, float tstd, double tmean_d,
                                                 double tstd_d, double eps_d, float *__restrict__ out)
{
#pragma clang fp contract(off)
    extern __shared__ float tcache[];
    const int x = blockIdx.x * 64 + (threadIdx.x & 63);
    const int y = blockIdx.y * 4 + (threadIdx.x >> 6);
    const int z = blockIdx.z;
    if (x >= w || y >= h) return;
    const int64_t o = ((int64_t)z * h + y) * w + x;
    if (x < brw || x >= w - brw || y < brh || y >= h - brh || z < brd || z >= d - brd) {
        out[o] = 0.0f;
        return;
    }
    const float cs = stds[o];
    if (M == SRED_ABSDIFF) {  // kernelGetBlockAbsDiffStds3D.cl:28, then 1 - d
        out[o] = 1.0f - fabsf(tstd - cs);
        return;
    }
    const float *c = img + o;
    const float cm = means[o];
    float r;
    if (M == SRED_PEARSON) {  // kernelGetBlockPearson3D.cl:58-112
        float mn = FLT_MAX, mx = -FLT_MAX;
        for (int k = 0; k < K; k++) {
            const float v = c[offs[k]];
            mn = fminf(mn, v);
            mx = fmaxf(mx, v);
        }
        const double den1 = (double)(mx - mn) + eps_d;
        float mn2 = FLT_MAX, mx2 = -FLT_MAX;
        for (int k = 0; k < K; k++) {
            const float t = (float)((double)(c[offs[k]] - mn) / den1) - cm;
            if (CACHE) tcache[k * 256 + threadIdx.x] = t;
            mn2 = fminf(mn2, t);
            mx2 = fmaxf(mx2, t);
        }
        const double den2 = (double)(mx2 - mn2) + eps_d;
        float cov = 0.0f;
        for (int k = 0; k < K; k++) {
            const float t = CACHE ? tcache[k * 256 + threadIdx.x] : (float)((double)(c[offs[k]] - mn) / den1) - cm;
            const float t3 = (float)((double)(t - mn2) / den2);
            cov += tmpl[k] * t3;
        }
        cov /= (float)(K - 1);
        if (tstd_d == 0.0 && cs == 0.0f) r = 1.0f;
        else if (tstd_d == 0.0 || cs == 0.0f) r = 0.0f;
        else r = fmaxf(0.0f, (float)((double)cov / (tstd_d * (double)cs + eps_d)));
    } else if (M == SRED_SSIM) {  // kernelGetBlockSsim3D.cl:66-103
        float cov = 0.0f;
        for (int k = 0; k < K; k++) cov += tmpl[k] * (c[offs[k]] - cm);
        cov /= (float)(K - 1);
        const float c1 = 0.0001f, c2 = 0.0009f;
        if (tstd_d == 0.0 && cs == 0.0f) r = 1.0f;
        else if (tstd_d == 0.0 || cs == 0.0f) r = 0.0f;
        else {
            const float num = (float)((2.0 * tmean_d * (double)cm + (double)c1) * (double)(2.0f * cov + c2));
            const double den =
                ((tmean_d * tmean_d) + (double)(cm * cm) + (double)c1) * ((tstd_d * tstd_d) + (double)(cs * cs) + (double)c2);
            r = fmaxf(0.0f, (float)((double)num / den));
        }
    } else {  // kernelGetBlockNrmse3D.cl:61-75, then 1/rmse
        float mse = 0.0f;
        for (int k = 0; k < K; k++) {
            const float e = tmpl[k] - (c[offs[k]] - cm);
            mse += e * e;
        }
        mse /= (float)K;
        const float rm = sqrtf(mse);
        r = (rm == 0.0f) ? 1.0f : 1.0f / rm;
    }
    out[o] = r;
}

}  // namespace

int block_map3d_dev(sred_ctx *ctx, const Geometry &g, int metric, const float *d_img, const float *h_tmpl,
                    float tmpl_mean, float tmpl_std, float c, int normalize, float *d_out, float *d_means,
                    float *d_stds, float *d_mask)
{
    if (metric < SRED_PEARSON || metric > SRED_ABSDIFF) {
        set_error("unknown metric");
        return SRED_ERR_UNKNOWN_METRIC;
    }
    // statistics (with the ellipsoid's volume offsets in ctx->offs3d)
    SRED_TRY(local_stats3d_dev(ctx, g, d_img, d_means, d_stds));
    if (ctx->host_tmpl3d.size() != (size_t)g.K || !std::equal(h_tmpl, h_tmpl + g.K, ctx->host_tmpl3d.begin()) ||
        !ctx->tmpl3d.ptr) {
        ctx->host_tmpl3d.clear();
        SRED_TRY(ctx->tmpl3d.ensure(sizeof(float) * (size_t)g.K));
        SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        SRED_HIP_CHECK(hipMemcpy(ctx->tmpl3d.ptr, h_tmpl, sizeof(float) * (size_t)g.K, hipMemcpyHostToDevice));
        ctx->host_tmpl3d.assign(h_tmpl, h_tmpl + g.K);
    }
    const double tm = java_float_literal(tmpl_mean), ts = java_float_literal(tmpl_std), ed = java_float_literal(kEps);
    const dim3 grid((g.w + 63) / 64, (g.h + 3) / 4, g.d);
    const int *offs = ctx->offs3d.as<int>();
    const float *tp = ctx->tmpl3d.as<float>();
    SRED_TRY(time_begin(ctx));
    switch (metric) {
    case SRED_PEARSON:
        if (g.K <= kCacheK)
            hipLaunchKernelGGL((k_block3d<SRED_PEARSON, true>), grid, dim3(256), sizeof(float) * 256 * (size_t)g.K,
                               ctx->stream, d_img, d_means, d_stds, g.w, g.h, g.d, g.brw, g.brh, g.brd, offs, tp, g.K,
                               tmpl_std, tm, ts, ed, d_out);
        else
            hipLaunchKernelGGL((k_block3d<SRED_PEARSON, false>), grid, dim3(256), 0, ctx->stream, d_img, d_means,
                               d_stds, g.w, g.h, g.d, g.brw, g.brh, g.brd, offs, tp, g.K, tmpl_std, tm, ts, ed, d_out);
        break;
    case SRED_SSIM:
        hipLaunchKernelGGL((k_block3d<SRED_SSIM, false>), grid, dim3(256), 0, ctx->stream, d_img, d_means, d_stds, g.w,
                           g.h, g.d, g.brw, g.brh, g.brd, offs, tp, g.K, tmpl_std, tm, ts, ed, d_out);
        break;
    case SRED_NRMSE:
        hipLaunchKernelGGL((k_block3d<SRED_NRMSE, false>), grid, dim3(256), 0, ctx->stream, d_img, d_means, d_stds,
                           g.w, g.h, g.d, g.brw, g.brh, g.brd, offs, tp, g.K, tmpl_std, tm, ts, ed, d_out);
        break;
    default:
        hipLaunchKernelGGL((k_block3d<SRED_ABSDIFF, false>), grid, dim3(256), 0, ctx->stream, d_img, d_means, d_stds,
                           g.w, g.h, g.d, g.brw, g.brh, g.brd, offs, tp, g.K, tmpl_std, tm, ts, ed, d_out);
        break;
    }
    SRED_HIP_CHECK(hipGetLastError());
    SRED_TRY(time_end(ctx, g.n()));
    // the 3D mask (Utils.getRelevanceMask3D, Utils.applyMask3D) when c > 0,
    // then Utils.normalizeImage3D over it (all inner voxels when c <= 0)
    const float *mask = nullptr;
    if (c > 0.0f) {
        SRED_TRY(vol_relevance_mask_dev(ctx, g, d_stds, c, d_mask, nullptr, nullptr, nullptr));
        SRED_TRY(apply_mask_dev(ctx, g, d_out, d_mask));
        mask = d_mask;
    }
    if (normalize) SRED_TRY(normalize_map_dev(ctx, g, d_out, mask, 0));
    return SRED_OK;
}

}  // namespace sred
