// vol.hip -- global repetition of volumes: GlobalRepetition3D_ with its four
// metrics (GlobalRepetition3D_.java:101-160, CLUtils.getLocalStatistics3D
// :1030-1114, calculateGlobalRepetitionMap3D :2793-2815 -> getGlobalPearson3D
// :2225-2335, getGlobalSsim3D :2362-2471, getGlobalNrmse3D :2496-2622,
// getGlobalAbsDiffStds3D :2647-2773, kernelGetLocalStatistics3D.cl,
// kernelGetGlobal{Pearson,Ssim,Nrmse,AbsDiffStds}3D.cl, Utils.getRelevanceMask3D
// :1501-1540).
//
// The reference's 3D path does not run as written; the intended computation
// is built, with four decisions:
//   * GlobalRepetition3D_ passes (radii..., blockSize) where
//     getLocalStatistics3D expects (blockSize, radii...): the statistics use
//     the block size and radii as meant;
//   * it builds the relevance mask with the 2D Utils.getRelevanceMask (a
//     w*h array, then indexed at w*h*z + ... for z >= 1): the 3D mask of
//     Utils.getRelevanceMask3D is used, whose noise variance is the float
//     mean of std^2 over the inner voxels, threshold noiseVar * c, mask 0
//     where std^2 <= threshold;
//   * the kernels gather the reference block over z < gz + bRZ but every test
//     block over z <= z + bRZ: both use the full ellipsoid;
//   * kernelGetGlobalSsim3D.cl:76 adds 0 to an undeclared pearson_map, so the
//     kernel does not compile: the no-op is dropped.
// Per pair the kernels are the 2D ones (weight, clamp, flat cases), except
// that a flat reference block scores 0 (it is skipped with the others below
// the threshold), NRMSE adds sqrt(mse) * weight instead of sqrt(mse)
// (kernelGetGlobalNrmse3D.cl:111) and AbsDiffStds also skips flat test
// blocks (kernelGetGlobalAbsDiffStds3D.cl:53), so flat blocks leave its
// compaction.  The pipeline is therefore the 2D one on a volume Geometry:
// compaction of valid voxels, ellipsoid offsets (dz*w*h + dy*w + dx) for the
// split-BF16 gather, the same symmetric MFMA contraction and weight sums, and
// the 3D interior for the finalize, the inversion and the normalisation.
#include <algorithm>
#include <cmath>
#include <vector>

#include "sred_internal.h"

namespace sred {

namespace {

constexpr int kVolSeg = 1024;  // elements per stage of the in-order chain (16 per lane)

// kernelGetLocalStatistics3D.cl: mean and unbiased std of the ellipsoid
// around every inner voxel, float sums in footprint order; borders 0.
// 64 x 4 voxels per workgroup (x across the wave: each offset is one
// coalesced row read, and no lanes idle on narrow volumes).
__global__ __launch_bounds__(256) void k_local_stats3d(const float *__restrict__ img, int w, int h, int d, int brw,
                                                       int brh, int brd, const int *__restrict__ offs, int K,
                                                       float *__restrict__ means, float *__restrict__ stds)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6), z = blockIdx.z;
    if (x >= w || y >= h) return;
    const int64_t o = ((int64_t)z * h + y) * w + x;
    if (x < brw || x >= w - brw || y < brh || y >= h - brh || z < brd || z >= d - brd) {
        means[o] = 0.0f;
        stds[o] = 0.0f;
        return;
    }
    const float *c = img + o;
    float m = 0.0f;
    for (int k = 0; k < K; k++) m += c[offs[k]];
    m /= (float)K;
    float v = 0.0f;
    for (int k = 0; k < K; k++) {
        const float t = c[offs[k]] - m;
        v += t * t;
    }
    means[o] = m;
    stds[o] = sqrtf(v / (float)(K - 1));
}

// Utils.getMeanNoiseVar3D(..., localStds) (:1463-1499): the float sum of
// std^2 over the inner voxels in raster order (one wave: the lanes load and
// square 1024 voxels, lane 0 adds them in order from LDS while the next
// 1024 loads are in flight), divided by the float count (which stops at 2^24).
__global__ __launch_bounds__(64) void k_vol_noise_var(const float *__restrict__ stds, int w, int h, int d, int brw,
                                                      int brh, int brd, float *__restrict__ out)
{
#pragma clang fp contract(off)
    __shared__ __align__(16) float sb[kVolSeg];
    const int iw = w - 2 * brw, ih = h - 2 * brh, id = d - 2 * brd;
    const int64_t n = (int64_t)iw * ih * id;
    const int lane = threadIdx.x;
    auto fetch = [&](int64_t e) {  // e < n < 2^31: 32-bit division
        if (e >= n) return 0.0f;
        const unsigned u = (unsigned)e, t = u / (unsigned)iw, z = t / (unsigned)ih;
        const int x = (int)(u - t * (unsigned)iw), y = (int)(t - z * (unsigned)ih);
        const float s = stds[((int64_t)(z + brd) * h + (y + brh)) * w + (x + brw)];
        return s * s;
    };
    float raw[kVolSeg / 64];
#pragma unroll
    for (int q = 0; q < kVolSeg / 64; q++) raw[q] = fetch(lane + 64 * q);
    float acc = 0.0f;
    for (int64_t base = 0; base < n; base += kVolSeg) {
#pragma unroll
        for (int q = 0; q < kVolSeg / 64; q++) sb[lane + 64 * q] = raw[q];
#pragma unroll
        for (int q = 0; q < kVolSeg / 64; q++) raw[q] = fetch(base + kVolSeg + lane + 64 * q);
        __syncthreads();
        if (lane == 0) {
            const int cnt = (int)std::min<int64_t>(kVolSeg, n - base);
            const float4 *s4 = reinterpret_cast<const float4 *>(sb);
            int g = 0;
#pragma unroll 8
            for (; g < (cnt >> 2); g++) {
                const float4 v = s4[g];
                acc += v.x;
                acc += v.y;
                acc += v.z;
                acc += v.w;
            }
            for (int t = g << 2; t < cnt; t++) acc += sb[t];
        }
        __syncthreads();
    }
    if (lane == 0) out[0] = acc / (float)std::min<int64_t>(n, 16777216);
}

// Utils.getRelevanceMask3D (:1501-1540): 0 where std^2 <= thr, 1 elsewhere
// inside the borders; count[0] += ones (GlobalRepetition3D_.java:134-145).
__global__ __launch_bounds__(256) void k_vol_mask(const float *__restrict__ stds, int w, int h, int d, int brw, int brh,
                                                  int brd, float thr, float *__restrict__ mask,
                                                  unsigned long long *__restrict__ count)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y, z = blockIdx.z;
    float m = 0.0f;
    if (x < w) {
        const int64_t o = ((int64_t)z * h + y) * w + x;
        if (x >= brw && x < w - brw && y >= brh && y < h - brh && z >= brd && z < d - brd) {
            const float s = stds[o];
            m = (s * s <= thr) ? 0.0f : 1.0f;
        }
        mask[o] = m;
    }
    const unsigned long long ones = __ballot(m > 0.0f);
    if ((threadIdx.x & 63) == 0 && ones) atomicAdd(count, (unsigned long long)__popcll(ones));
}

}  // namespace

int upload_offsets3d(sred_ctx *ctx, const Geometry &g)
{
    const int key[5] = {g.bw, g.bh, g.bd, g.w, g.h};
    if (ctx->offs3d.ptr && std::equal(key, key + 5, ctx->offs3d_key)) return SRED_OK;
    std::vector<int> dx, dy, dz;
    footprint3d(g.bw, g.bh, g.bd, &dx, &dy, &dz);
    std::vector<int> off(g.K);
    const int64_t plane = (int64_t)g.w * g.h;
    for (int k = 0; k < g.K; k++) off[k] = (int)(dz[k] * plane + (int64_t)dy[k] * g.w + dx[k]);
    SRED_TRY(ctx->offs3d.ensure(sizeof(int) * g.K));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    SRED_HIP_CHECK(hipMemcpy(ctx->offs3d.ptr, off.data(), sizeof(int) * g.K, hipMemcpyHostToDevice));
    std::copy(key, key + 5, ctx->offs3d_key);
    return SRED_OK;
}

int local_stats3d_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, float *d_means, float *d_stds)
{
    SRED_TRY(upload_offsets3d(ctx, g));
    hipLaunchKernelGGL(k_local_stats3d, dim3((g.w + 63) / 64, (g.h + 3) / 4, g.d), dim3(256), 0, ctx->stream, d_img,
                       g.w, g.h, g.d, g.brw, g.brh, g.brd, ctx->offs3d.as<int>(), g.K, d_means, d_stds);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

int vol_relevance_mask_dev(sred_ctx *ctx, const Geometry &g, const float *d_stds, float c, float *d_mask,
                           float *noise_var, float *threshold, float *n_pixels)
{
#pragma clang fp contract(off)
    SRED_TRY(ctx->misc.ensure(kMiscBytes));
    float *d_nv = reinterpret_cast<float *>(ctx->misc.as<char>() + 2048);
    unsigned long long *d_cnt = reinterpret_cast<unsigned long long *>(ctx->misc.as<char>() + 2048 + 64);
    const bool inner = g.w > 2 * g.brw && g.h > 2 * g.brh && g.d > 2 * g.brd;
    float nv = 0.0f;
    if (inner) {
        hipLaunchKernelGGL(k_vol_noise_var, dim3(1), dim3(64), 0, ctx->stream, d_stds, g.w, g.h, g.d, g.brw, g.brh,
                           g.brd, d_nv);
        SRED_HIP_CHECK(hipMemcpyAsync(&nv, d_nv, sizeof nv, hipMemcpyDeviceToHost, ctx->stream));
        SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    } else {
        nv = NAN;  // 0 / 0 in the reference
    }
    const float thr = nv * c;
    SRED_TRY(zero_dev(ctx, d_cnt, sizeof(unsigned long long)));
    hipLaunchKernelGGL(k_vol_mask, dim3((g.w + 255) / 256, g.h, g.d), dim3(256), 0, ctx->stream, d_stds, g.w, g.h, g.d,
                       g.brw, g.brh, g.brd, thr, d_mask, d_cnt);
    SRED_HIP_CHECK(hipGetLastError());
    unsigned long long cnt = 0;
    SRED_HIP_CHECK(hipMemcpyAsync(&cnt, d_cnt, sizeof cnt, hipMemcpyDeviceToHost, ctx->stream));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (noise_var) *noise_var = nv;
    if (threshold) *threshold = thr;
    if (n_pixels) *n_pixels = (float)std::min<unsigned long long>(cnt, 16777216ull);  // float += 1.0f stops at 2^24
    return SRED_OK;
}

int global_map3d_dev(sred_ctx *ctx, const Geometry &g, int metric, int precision, const float *d_img, float c,
                     int normalize, float *d_out, float *d_means, float *d_stds, float *d_mask, float *d_S,
                     float *d_W)
{
    SRED_TRY(local_stats3d_dev(ctx, g, d_img, d_means, d_stds));
    float nv, thr, np;
    SRED_TRY(vol_relevance_mask_dev(ctx, g, d_stds, c, d_mask, &nv, &thr, &np));
    GlobalParams p;
    p.metric = metric;
    p.precision = precision;
    p.filter_param = nv;
    p.threshold = thr;
    p.skip_flat_rows = true;
    p.nrmse_weighted = (metric == SRED_NRMSE);
    p.drop_flat = (metric == SRED_ABSDIFF);
    SRED_TRY(global_accum_dev(ctx, g, p, d_img, d_means, d_stds, d_S, d_W));
    return global_finalize_dev(ctx, g, metric, d_S, d_W, d_mask, np, c, normalize, d_out);
}

}  // namespace sred
