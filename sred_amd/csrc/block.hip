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
#include <vector>

#include "sred_internal.h"

namespace sred {

// Workgroup: 16 x 16 threads, each computing 4 horizontally adjacent windows,
// so one workgroup covers 64 x 16 output pixels.
constexpr int kBx = 64, kBy = 16;
constexpr int kMaxRows = SRED_MAX_BLOCK_SIDE + 1;

// Ellipse rows: half-width hw[j] and first footprint index kb[j] of row j
// (dy = j - brh); the footprint is row-major and every row is the contiguous
// run dx = -hw..hw (the float predicate is monotone in |dx|).
struct RowTab {
    int hw[kMaxRows];
    int kb[kMaxRows];
};

// One ellipse row of the four windows of a thread, as a sliding run: element
// e (0 <= e < L + 3) of the union of the four runs is loaded once (load(e))
// and is element e - w of window w's run when 0 <= e - w < L (f(w)).  Every
// window sees its elements in its own footprint order, so each sum keeps the
// reference's sequential float order.  L is odd: a single-element row gives
// window w element e = w; otherwise (L >= 3) the first and last three
// elements have fixed window sets and the rest feed all four windows.
template <class LD, class F>
__device__ __forceinline__ void row_run(int L, LD &&load, F &&f)
{
    if (L == 1) {
#pragma unroll
        for (int w = 0; w < 4; w++) {
            load(w);
            f(w);
        }
        return;
    }
    load(0);
    f(0);
    load(1);
    f(0);
    f(1);
    load(2);
    f(0);
    f(1);
    f(2);
    for (int e = 3; e < L; e++) {
        load(e);
#pragma unroll
        for (int w = 0; w < 4; w++) f(w);
    }
    load(L);
    f(1);
    f(2);
    f(3);
    load(L + 1);
    f(2);
    f(3);
    load(L + 2);
    f(3);
}

// LDS tile: row stride RS = 1 mod 64, so the 64 lanes of a wave (16 column
// groups 4 words apart, 4 rows) address 64 distinct banks.
// tp4[i] = (t[i], t[i-1], t[i-2], t[i-3]) (t = template, 0 outside 0..K-1):
// element e of row j is template entry kb[j] + e - w for window w.
template <int M>
__global__ __launch_bounds__(256) void k_block(const float *__restrict__ img, int w, int h, int brw, int brh, int K,
                                               RowTab rt, const float4 *__restrict__ tp4, double tmpl_mean_d,
                                               double tmpl_std_d, float tmpl_std_f, double eps_d,
                                               float *__restrict__ out, float *__restrict__ means_out,
                                               float *__restrict__ stds_out, unsigned int *__restrict__ mmpart)
{
#pragma clang fp contract(off)
    extern __shared__ float sm[];
    const int tw = kBx + 2 * brw, th = kBy + 2 * brh;
    const int RS = (tw + 62) / 64 * 64 + 1;
    float *tile = sm;
    float4 *tp = reinterpret_cast<float4 *>(sm + ((th * RS + 3) & ~3));
    const int x0 = blockIdx.x * kBx, y0 = blockIdx.y * kBy;
    if (blockIdx.z) {  // frame of a stack (batched frames: no statistics outputs)
        const int64_t zoff = (int64_t)blockIdx.z * w * h;
        img += zoff;
        out += zoff;
    }
    for (int i = threadIdx.x; i < tw * th; i += blockDim.x) {
        const int ly = i / tw, lx = i - ly * tw;
        const int gy = y0 - brh + ly, gx = x0 - brw + lx;
        tile[ly * RS + lx] =
            (gx >= 0 && gx < w && gy >= 0 && gy < h) ? img[(int64_t)gy * w + gx] : 0.0f;
    }
    if (M != SRED_ABSDIFF)
        for (int i = threadIdx.x; i < K + 3; i += blockDim.x) tp[i] = tp4[i];
    __syncthreads();
    const int tx = threadIdx.x & 15, ty = threadIdx.x >> 4;
    const int gx0 = x0 + 4 * tx, gy = y0 + ty;
    unsigned int lo = 0xffffffffu, hi = 0u;  // interior min/max of this thread's outputs (ordered ints)
    if (gx0 < w && gy < h) {
        const int nrows = 2 * brh + 1;
        const float *trow = tile + ty * RS + 4 * tx;  // the windows' top row, this thread's first window

        // window means, kernelGetLocalStatistics2D.cl:40-47
        float m[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int j = 0; j < nrows; j++) {
            const int hw = rt.hw[j], c0 = brw - hw;
            const float *rp = trow + j * RS;
            float v = 0.0f;
            row_run(
                2 * hw + 1,
                [&](int e) { v = rp[c0 + e]; },
                [&](int wi) { m[wi] += v; });
        }
    #pragma unroll
        for (int wi = 0; wi < 4; wi++) m[wi] /= (float)K;

        // variance (kernelGetLocalStatistics2D.cl:48-55) and the template sum
        // (covariance, or squared error for NRMSE) in one sweep: the two sums are
        // independent sequences, each in footprint order
        float var[4] = {0.0f, 0.0f, 0.0f, 0.0f}, acc[4] = {0.0f, 0.0f, 0.0f, 0.0f};
        for (int j = 0; j < nrows; j++) {
            const int hw = rt.hw[j], c0 = brw - hw, kb = rt.kb[j];
            const float *rp = trow + j * RS;
            float v = 0.0f;
            float4 t = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            row_run(
                2 * hw + 1,
                [&](int e) {
                    v = rp[c0 + e];
                    if (M != SRED_ABSDIFF) t = tp[kb + e];
                },
                [&](int wi) {
                    const float d = v - m[wi];
                    var[wi] += d * d;
                    if (M == SRED_PEARSON || M == SRED_SSIM) {
                        const float tv = wi == 0 ? t.x : wi == 1 ? t.y : wi == 2 ? t.z : t.w;
                        acc[wi] += tv * d;  // kernelGetBlockPearson2D.cl:70-74
                    } else if (M == SRED_NRMSE) {
                        const float tv = wi == 0 ? t.x : wi == 1 ? t.y : wi == 2 ? t.z : t.w;
                        const float q = tv - d;  // kernelGetBlockNrmse2D.cl:69-72
                        acc[wi] += q * q;
                    }
                });
        }

    #pragma unroll
        for (int wi = 0; wi < 4; wi++) {
            const int gx = gx0 + wi;
            if (gx >= w) break;
            const int64_t o = (int64_t)gy * w + gx;
            if (gx < brw || gx >= w - brw || gy < brh || gy >= h - brh) {
                out[o] = 0.0f;
                if (means_out) { means_out[o] = 0.0f; stds_out[o] = 0.0f; }
                continue;
            }
            const float cs = sqrtf(var[wi] / (float)(K - 1));
            if (means_out) { means_out[o] = m[wi]; stds_out[o] = cs; }
            float r;
            if (M == SRED_ABSDIFF) {  // kernelGetBlockAbsDiffStds2D.cl:103-105
                r = fabsf(tmpl_std_f - cs);
            } else if (M == SRED_NRMSE) {  // kernelGetBlockNrmse2D.cl:73-74
                r = sqrtf(acc[wi] / (float)K);
            } else {
                const float cov = acc[wi] / (float)(K - 1);
                if (tmpl_std_d == 0.0 && cs == 0.0f) {
                    r = 1.0f;
                } else if (tmpl_std_d == 0.0 || cs == 0.0f) {
                    r = 0.0f;
                } else if (M == SRED_PEARSON) {  // kernelGetBlockPearson2D.cl:88 (double literals)
                    r = fmaxf(0.0f, (float)((double)cov / ((tmpl_std_d * (double)cs) + eps_d)));
                } else {  // kernelGetBlockSsim2D.cl:169 (sum denominator, double literals)
                    const float c1 = 0.0001f, c2 = 0.0009f;
                    const float mm = m[wi];
                    double num = (2.0 * tmpl_mean_d * (double)mm + (double)c1) * (double)(2.0f * cov + c2);
                    double den = ((tmpl_mean_d * tmpl_mean_d) + (double)(mm * mm) + (double)c1) *
                                 ((tmpl_std_d * tmpl_std_d) + (double)(cs * cs) + (double)c2);
                    r = fmaxf(0.0f, (float)(num / den));
                }
            }
            out[o] = r;
            lo = min(lo, f2ord(r));
            hi = max(hi, f2ord(r));
        }
    }
    if (mmpart) {  // workgroup partial of the normalisation's min/max (Utils.java:991-1012)
        __shared__ unsigned int slo[256], shi[256];
        slo[threadIdx.x] = lo;
        shi[threadIdx.x] = hi;
        __syncthreads();
        for (int s = 128; s > 0; s >>= 1) {
            if (threadIdx.x < s) {
                slo[threadIdx.x] = min(slo[threadIdx.x], slo[threadIdx.x + s]);
                shi[threadIdx.x] = max(shi[threadIdx.x], shi[threadIdx.x + s]);
            }
            __syncthreads();
        }
        if (threadIdx.x == 0) {
            const int b = (blockIdx.z * gridDim.y + blockIdx.y) * gridDim.x + blockIdx.x;
            mmpart[3 * b] = slo[0];
            mmpart[3 * b + 1] = shi[0];
            mmpart[3 * b + 2] = 0u;
        }
    }
}

int64_t block_workgroups(const Geometry &g) { return (int64_t)((g.w + kBx - 1) / kBx) * ((g.h + kBy - 1) / kBy); }

int block_raw_dev(sred_ctx *ctx, const Geometry &g, int metric, const float *d_img, float *d_means_out,
                  float *d_stds_out, const float *h_tmpl, float tmpl_mean, float tmpl_std, float *d_out,
                  unsigned int *d_mmpart, int *nparts, int nframes)
{
    if (nframes < 1 || nframes > 65535 || (nframes > 1 && (d_means_out || d_stds_out))) {
        set_error("invalid frame count for the block stencil");
        return SRED_ERR_ARG;
    }
    if (metric < 0 || metric > 3) {
        set_error("unknown metric");
        return SRED_ERR_UNKNOWN_METRIC;
    }
    // ellipse rows as runs (row-major footprint)
    std::vector<int> dx, dy;
    footprint(g.bw, g.bh, &dx, &dy);
    RowTab rt{};
    int k = 0;
    for (int j = 0; j < g.bh; j++) {
        const int y = j - g.brh;
        rt.kb[j] = k;
        int n = 0;
        while (k + n < g.K && dy[k + n] == y) n++;
        if (n == 0 || n % 2 == 0 || dx[k] != -(n / 2) || dx[k + n - 1] != n / 2) {
            set_error("unsupported block footprint");
            return SRED_ERR_ARG;
        }
        rt.hw[j] = n / 2;
        k += n;
    }
    // expanded template (t[i], t[i-1], t[i-2], t[i-3])
    std::vector<float> tp4(4 * (size_t)(g.K + 3), 0.0f);
    for (int i = 0; i < g.K + 3; i++)
        for (int w = 0; w < 4; w++) {
            const int q = i - w;
            tp4[4 * i + w] = (q >= 0 && q < g.K) ? h_tmpl[q] : 0.0f;
        }
    if (tp4 != ctx->host_tmpl || !ctx->tmpl.ptr) {
        ctx->host_tmpl.clear();
        SRED_TRY(ctx->tmpl.ensure(sizeof(float) * tp4.size()));
        SRED_HIP_CHECK(hipMemcpy(ctx->tmpl.ptr, tp4.data(), sizeof(float) * tp4.size(), hipMemcpyHostToDevice));
        ctx->host_tmpl = tp4;
    }
    const int tw = kBx + 2 * g.brw, th = kBy + 2 * g.brh;
    const int RS = (tw + 62) / 64 * 64 + 1;
    const size_t lds = sizeof(float) * (size_t)((th * RS + 3) & ~3) + sizeof(float4) * (size_t)(g.K + 3);
    dim3 grid((g.w + kBx - 1) / kBx, (g.h + kBy - 1) / kBy, nframes);
    if (nparts) *nparts = (int)(grid.x * grid.y);
    float *mo = d_stds_out ? d_means_out : nullptr, *so = d_means_out ? d_stds_out : nullptr;
    const float4 *d_tp = ctx->tmpl.as<float4>();
    const double tm = java_float_literal(tmpl_mean), ts = java_float_literal(tmpl_std), ed = java_float_literal(kEps);
    SRED_TRY(time_begin(ctx));
    switch (metric) {
    case SRED_PEARSON:
        hipLaunchKernelGGL(k_block<SRED_PEARSON>, grid, dim3(256), lds, ctx->stream, d_img, g.w, g.h, g.brw, g.brh,
                           g.K, rt, d_tp, tm, ts, tmpl_std, ed, d_out, mo, so, d_mmpart);
        break;
    case SRED_SSIM:
        hipLaunchKernelGGL(k_block<SRED_SSIM>, grid, dim3(256), lds, ctx->stream, d_img, g.w, g.h, g.brw, g.brh, g.K,
                           rt, d_tp, tm, ts, tmpl_std, ed, d_out, mo, so, d_mmpart);
        break;
    case SRED_NRMSE:
        hipLaunchKernelGGL(k_block<SRED_NRMSE>, grid, dim3(256), lds, ctx->stream, d_img, g.w, g.h, g.brw, g.brh,
                           g.K, rt, d_tp, tm, ts, tmpl_std, ed, d_out, mo, so, d_mmpart);
        break;
    default:
        hipLaunchKernelGGL(k_block<SRED_ABSDIFF>, grid, dim3(256), lds, ctx->stream, d_img, g.w, g.h, g.brw, g.brh,
                           g.K, rt, d_tp, tm, ts, tmpl_std, ed, d_out, mo, so, d_mmpart);
    }
    SRED_HIP_CHECK(hipGetLastError());
    SRED_TRY(time_end(ctx, g.n() * nframes));
    return SRED_OK;
}

}  // namespace sred
