// adjoint.hip -- gradient of the global Pearson repetition map with respect to
// the (normalised) input pixels, by a hand-written adjoint.
//
// Forward (kernelGetGlobalPearson2D.cl:67-166, CLUtils.java:1747-1770 with
// relevance constant 0 and normalisation on):
//   T_ij = max(0, r_ij) w_ij,  r_ij = cov_ij / (s_i s_j + eps),
//   w_ij = exp(-(s_i - s_j)^2 / H),  H = h + eps,
//   S_i = sum_j T_ij,  W_i = sum_j w_ij,  rep_i = S_i / (W_i n + eps),
//   map_i = (rep_i - mn) / (mx - mn + eps) where the mask is 1, rep_i else.
// Held fixed (stop-gradient): h, n (nPixels), the mask and mn/mx; blocks that
// are flat (s == 0) keep zero statistics (no gradient through them).
//
// With g_i = dL/dmap_i * dmap_i/drep_i, D_i = W_i n + eps,
// a_i = g_i / D_i (0 for flat rows, whose S_i is constant) and
// b_i = -g_i rep_i n / D_i, every pair enters the loss through
// a_i dT_ij + b_i dw_ij.  Because T and w are symmetric in (i, j), the pair
// (j, i) contributes the same terms with a_j, b_j, so one row-stationary pass
// gives the full gradient of row i's block:
//   M_ij   = [r_ij > 0] w_ij / (s_i s_j + eps)
//   dL/dA_i  = sum_j (a_i + a_j) M_ij A_j / (K-1)                    (block vector)
//   dL/ds_i  = sum_j -(a_i + a_j) M_ij r_ij s_j
//              + w_ij (-2 (s_i - s_j) / H) ((a_i + a_j) [r>0] r_ij + b_i + b_j)
// then ds_i/dA_i = A_i / ((K-1) s_i), the mean subtraction projects out the
// block mean, and a gather-transpose stencil adds each block's K entries back
// onto the pixels they came from.  Sums run in FP32 within a 64-column tile
// and FP64 across tiles.
#include <float.h>

#include <algorithm>
#include <cmath>

#include "sred_internal.h"

namespace sred {

// Per valid block: (s, a, b, t) with t = s*sqrt(log2 e / H) for the weight.
__global__ __launch_bounds__(256) void k_adjoint_coef(const int *__restrict__ idx, int Nv, int Nv_pad,
                                                      const float *__restrict__ stds, const float *__restrict__ S,
                                                      const float *__restrict__ W, const float *__restrict__ mask,
                                                      const float *__restrict__ dmap, const unsigned int *__restrict__ mm,
                                                      float n_pixels, float tscale, float4 *__restrict__ info)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Nv_pad) return;
    if (v >= Nv) {
        info[v] = make_float4(0.0f, 0.0f, 0.0f, 1e30f);  // padding: weight exp2(-inf) = 0, a = b = 0
        return;
    }
    int p = idx[v];
    // normalisation constants of the forward map (ordered-int min/max)
    unsigned int lo = mm[0], hi = mm[1];
    float mn = __uint_as_float((lo & 0x80000000u) ? (lo & 0x7fffffffu) : ~lo);
    float mx = __uint_as_float((hi & 0x80000000u) ? (hi & 0x7fffffffu) : ~hi);
    double g = (double)dmap[p];
    if (mask[p] > 0.0f) g /= ((double)mx - (double)mn + 1e-7);
    double D = (double)W[p] * (double)n_pixels + 1e-7;
    double rep = (double)S[p] / D;
    float s = stds[p];
    double a = (s == 0.0f) ? 0.0 : g / D;
    double b = -g * rep * (double)n_pixels / D;
    info[v] = make_float4(s, (float)a, (float)b, s * tscale);
}

__global__ void k_fill_int(int *__restrict__ p, int64_t n, int value)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = value;
}

__global__ void k_pix2v(const int *__restrict__ idx, int Nv, int *__restrict__ pix2v)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v < Nv) pix2v[idx[v]] = v;
}

// Row-stationary adjoint pass: 64 rows per workgroup against every column
// tile.  UT is the Pearson-scaled block matrix u = (x - mu)/(s sqrt(K-1)),
// k-major [K4][Nv_pad]; Gc receives, per row, the centred gradient of its K
// block pixels ([Nv_pad][K16], zero for flat and padding rows).
template <int M16>
__global__ __launch_bounds__(256) void k_adjoint_f32(const float *__restrict__ UT, const float4 *__restrict__ info,
                                                     int K, int K4, int Nv, int Nv_pad, float eps, float m2_over_H,
                                                     float inv_sqrt_km1, float *__restrict__ Gc)
{
    constexpr int K16 = 16 * M16;            // k values: tx + 16 m, m < M16
    extern __shared__ float sm[];
    float *As = sm;                          // [K4][64]
    float *Bs = As + (size_t)K4 * 64;        // [K4][64]
    float *Bt = Bs + (size_t)K4 * 64;        // [64][K16 + 1]
    float *Pt = Bt + (size_t)64 * (K16 + 1); // [64][65]
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int row0 = blockIdx.x * 64;

    for (int i = tid; i < K4 * 16; i += 256) {
        int k = i >> 4, q = i & 15;
        *reinterpret_cast<float4 *>(&As[k * 64 + q * 4]) =
            *reinterpret_cast<const float4 *>(&UT[(size_t)k * Nv_pad + row0 + q * 4]);
    }
    float4 Ri[4];
#pragma unroll
    for (int i = 0; i < 4; i++) Ri[i] = info[row0 + ty * 4 + i];

    double gs[4] = {0, 0, 0, 0};
    double ga[4][M16];
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int m = 0; m < M16; m++) ga[i][m] = 0.0;

    for (int c0 = 0; c0 < Nv_pad; c0 += 64) {
        __syncthreads();
        for (int i = tid; i < K4 * 16; i += 256) {
            int k = i >> 4, q = i & 15;
            float4 v = *reinterpret_cast<const float4 *>(&UT[(size_t)k * Nv_pad + c0 + q * 4]);
            *reinterpret_cast<float4 *>(&Bs[k * 64 + q * 4]) = v;
            Bt[(q * 4 + 0) * (K16 + 1) + k] = v.x;
            Bt[(q * 4 + 1) * (K16 + 1) + k] = v.y;
            Bt[(q * 4 + 2) * (K16 + 1) + k] = v.z;
            Bt[(q * 4 + 3) * (K16 + 1) + k] = v.w;
        }
        for (int i = tid; i < 64 * (K16 - K4); i += 256) {  // zero k padding of Bt
            int j = i / (K16 - K4), k = K4 + i % (K16 - K4);
            Bt[j * (K16 + 1) + k] = 0.0f;
        }
        __syncthreads();
        float d[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) d[i][j] = 0.0f;
        for (int k = 0; k < K4; k++) {
            float4 a = *reinterpret_cast<const float4 *>(&As[k * 64 + ty * 4]);
            float4 b = *reinterpret_cast<const float4 *>(&Bs[k * 64 + tx * 4]);
            float av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
            for (int i = 0; i < 4; i++)
#pragma unroll
                for (int j = 0; j < 4; j++) d[i][j] = fmaf(av[i], bv[j], d[i][j]);
        }
        float4 Cj[4];
#pragma unroll
        for (int j = 0; j < 4; j++) Cj[j] = info[c0 + tx * 4 + j];
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float gsi = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; j++) {
                const float si = Ri[i].x, sj = Cj[j].x;
                const float dt = Ri[i].w - Cj[j].w;
                const float w = __builtin_amdgcn_exp2f(-(dt * dt));
                const float A = Ri[i].y + Cj[j].y;
                float r = 0.0f, M = 0.0f;
                if (si != 0.0f && sj != 0.0f) {
                    const float ss = si * sj;
                    const float den = ss + eps;
                    r = d[i][j] * ss / den;      // cov / (s_i s_j + eps)
                    if (r > 0.0f) M = w / den;
                    else r = 0.0f;
                }
                Pt[(ty * 4 + i) * 65 + tx * 4 + j] = A * M * sj;
                gsi += -A * M * r * sj + w * m2_over_H * (si - sj) * (A * r + Ri[i].z + Cj[j].z);
            }
            gs[i] += (double)gsi;
        }
        __syncthreads();
        float gt[4][M16];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int m = 0; m < M16; m++) gt[i][m] = 0.0f;
        for (int j = 0; j < 64; j++) {
            float bj[M16];
#pragma unroll
            for (int m = 0; m < M16; m++) bj[m] = Bt[j * (K16 + 1) + tx + 16 * m];
#pragma unroll
            for (int i = 0; i < 4; i++) {
                const float pij = Pt[(ty * 4 + i) * 65 + j];
#pragma unroll
                for (int m = 0; m < M16; m++) gt[i][m] = fmaf(pij, bj[m], gt[i][m]);
            }
        }
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int m = 0; m < M16; m++) ga[i][m] += (double)gt[i][m];
    }
    // dL/ds_i: reduce over the 16 column lanes
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) gs[i] += __shfl_xor(gs[i], m);
    // per row: G_k = (sum_j P''_ij u_jk + dL/ds_i u_ik) / sqrt(K-1), centred over k < K
#pragma unroll
    for (int i = 0; i < 4; i++) {
        const int v = row0 + ty * 4 + i;
        const bool live = v < Nv && Ri[i].x != 0.0f;
        double gk[M16], part = 0.0;
#pragma unroll
        for (int m = 0; m < M16; m++) {
            int k = tx + 16 * m;
            double uik = (k < K4) ? (double)As[k * 64 + ty * 4 + i] : 0.0;
            gk[m] = (k < K) ? (ga[i][m] + gs[i] * uik) * (double)inv_sqrt_km1 : 0.0;
            part += gk[m];
        }
#pragma unroll
        for (int msk = 1; msk < 16; msk <<= 1) part += __shfl_xor(part, msk);
        const double mean = part / (double)K;
#pragma unroll
        for (int m = 0; m < M16; m++) {
            int k = tx + 16 * m;
            Gc[(size_t)v * K16 + k] = (live && k < K) ? (float)(gk[m] - mean) : 0.0f;
        }
    }
}

// dL/dx_q = sum_k Gc[block centred at q - off_k][k] (transpose of the gather).
__global__ __launch_bounds__(256) void k_adjoint_scatter(const float *__restrict__ Gc, const int *__restrict__ pix2v,
                                                         const int *__restrict__ dxy, int K, int K16, int w, int h,
                                                         int brw, int brh, float *__restrict__ grad)
{
    int64_t n = (int64_t)w * h;
    for (int64_t q = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; q < n; q += (int64_t)gridDim.x * blockDim.x) {
        int x = (int)(q % w), y = (int)(q / w);
        double acc = 0.0;
        for (int k = 0; k < K; k++) {
            int cx = x - dxy[2 * k], cy = y - dxy[2 * k + 1];
            if (cx < brw || cx >= w - brw || cy < brh || cy >= h - brh) continue;
            int v = pix2v[(int64_t)cy * w + cx];
            if (v >= 0) acc += (double)Gc[(size_t)v * K16 + k];
        }
        grad[q] = (float)acc;
    }
}

template <int M16>
static void launch_adjoint_m(dim3 grid, size_t lds, hipStream_t st, const float *UT, const float4 *info, int K, int K4,
                             int Nv, int Nv_pad, float eps, float m2H, float isk, float *Gc)
{
    hipLaunchKernelGGL((k_adjoint_f32<M16>), grid, dim3(256), lds, st, UT, info, K, K4, Nv, Nv_pad, eps, m2H, isk, Gc);
}

static int launch_adjoint(int M16, dim3 grid, size_t lds, hipStream_t st, const float *UT, const float4 *info, int K,
                          int K4, int Nv, int Nv_pad, float eps, float m2H, float isk, float *Gc)
{
    switch (M16) {
    case 1: launch_adjoint_m<1>(grid, lds, st, UT, info, K, K4, Nv, Nv_pad, eps, m2H, isk, Gc); break;
    case 2: launch_adjoint_m<2>(grid, lds, st, UT, info, K, K4, Nv, Nv_pad, eps, m2H, isk, Gc); break;
    case 3: launch_adjoint_m<3>(grid, lds, st, UT, info, K, K4, Nv, Nv_pad, eps, m2H, isk, Gc); break;
    case 4: launch_adjoint_m<4>(grid, lds, st, UT, info, K, K4, Nv, Nv_pad, eps, m2H, isk, Gc); break;
    case 5: launch_adjoint_m<5>(grid, lds, st, UT, info, K, K4, Nv, Nv_pad, eps, m2H, isk, Gc); break;
    case 6: launch_adjoint_m<6>(grid, lds, st, UT, info, K, K4, Nv, Nv_pad, eps, m2H, isk, Gc); break;
    case 7: launch_adjoint_m<7>(grid, lds, st, UT, info, K, K4, Nv, Nv_pad, eps, m2H, isk, Gc); break;
    case 8: launch_adjoint_m<8>(grid, lds, st, UT, info, K, K4, Nv, Nv_pad, eps, m2H, isk, Gc); break;
    case 9: launch_adjoint_m<9>(grid, lds, st, UT, info, K, K4, Nv, Nv_pad, eps, m2H, isk, Gc); break;
    case 10: launch_adjoint_m<10>(grid, lds, st, UT, info, K, K4, Nv, Nv_pad, eps, m2H, isk, Gc); break;
    default:
        set_error("the adjoint supports blocks up to K = 160 pixels");
        return SRED_ERR_ARG;
    }
    return SRED_OK;
}

int global_pearson_adjoint_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, const float *d_dmap,
                               float *d_grad)
{
    const int64_t n = g.n();
    const int K = g.K, K4 = (K + 3) / 4 * 4, K16 = (K + 15) / 16 * 16;
    if (K16 / 16 > 10) {
        set_error("the adjoint supports blocks up to K = 160 pixels");
        return SRED_ERR_ARG;
    }
    const size_t lds = sizeof(float) * ((size_t)2 * K4 * 64 + (size_t)64 * (K16 + 1) + 64 * 65);
    if (lds > 160 * 1024) {
        set_error("block too large for the adjoint tile");
        return SRED_ERR_ARG;
    }
    // forward: statistics, noise variance, mask, accumulators (FP32) and the normalised map
    SRED_TRY(ctx->means.ensure(sizeof(float) * n));
    SRED_TRY(ctx->stds.ensure(sizeof(float) * n));
    SRED_TRY(ctx->mask.ensure(sizeof(float) * n));
    SRED_TRY(ctx->S.ensure(sizeof(float) * n));
    SRED_TRY(ctx->W.ensure(sizeof(float) * n));
    SRED_TRY(ctx->out.ensure(sizeof(float) * n));
    float *d_means = ctx->means.as<float>(), *d_stds = ctx->stds.as<float>(), *d_mask = ctx->mask.as<float>();
    float *d_S = ctx->S.as<float>(), *d_W = ctx->W.as<float>(), *d_map = ctx->out.as<float>();
    SRED_TRY(launch_local_stats(ctx, g, d_img, d_means, d_stds));
    float nv, np;
    SRED_TRY(noise_var_dev(ctx, d_img, g.w, g.h, &nv));
    SRED_TRY(relevance_mask_dev(ctx, g, d_stds, 0.0f, d_mask, &np));
    GlobalParams p;
    p.metric = SRED_PEARSON;
    p.precision = SRED_FP32;
    p.filter_param = nv;
    p.threshold = 0.0f;
    SRED_TRY(global_accum_dev(ctx, g, p, d_img, d_means, d_stds, d_S, d_W));
    SRED_TRY(global_finalize_dev(ctx, g, SRED_PEARSON, d_S, d_W, d_mask, np, 0.0f, 1, d_map));
    const unsigned int *d_mm = reinterpret_cast<const unsigned int *>(ctx->misc.as<char>() + 1024);
    const int Nv = ctx->last_nv;
    SRED_TRY(zero_dev(ctx, d_grad, sizeof(float) * n));
    if (Nv == 0) return SRED_OK;
    const int Nv_pad = (Nv + 63) / 64 * 64;
    const int *d_idx = ctx->valid_idx.as<int>();

    const double H = java_float_literal(nv) + java_float_literal(kEps);
    const float tscale = (float)std::sqrt(1.4426950408889634 / H);
    const float inv_sqrt_km1 = (float)(1.0 / std::sqrt((double)(K - 1)));
    SRED_TRY(ctx->aux.ensure(sizeof(float4) * (size_t)Nv_pad));
    float4 *d_info = ctx->aux.as<float4>();
    hipLaunchKernelGGL(k_adjoint_coef, dim3((Nv_pad + 255) / 256), dim3(256), 0, ctx->stream, d_idx, Nv, Nv_pad,
                       d_stds, d_S, d_W, d_mask, d_dmap, d_mm, np, tscale, d_info);
    // Pearson-scaled block vectors, k-major
    SRED_TRY(upload_offsets(ctx, g, g.w, &ctx->offs2, ctx->offs2_key));
    size_t nu = (size_t)Nv_pad * K4;
    SRED_TRY(ctx->u32.ensure(sizeof(float) * nu));
    launch_gather_f32(ctx, SRED_PEARSON, d_img, d_means, d_stds, d_idx, Nv, Nv_pad, K, K4, ctx->offs2.as<int>(),
                      inv_sqrt_km1, (float)(1.0 / std::sqrt((double)K)), ctx->u32.as<float>());
    SRED_TRY(ctx->partial.ensure(sizeof(float) * (size_t)Nv_pad * K16));
    float *d_Gc = ctx->partial.as<float>();
    SRED_TRY(time_begin(ctx));
    SRED_TRY(launch_adjoint(K16 / 16, dim3(Nv_pad / 64), lds, ctx->stream, ctx->u32.as<float>(), d_info, K, K4, Nv,
                            Nv_pad, kEps, (float)(-2.0 / H), inv_sqrt_km1, d_Gc));
    SRED_HIP_CHECK(hipGetLastError());
    SRED_TRY(time_end(ctx, (int64_t)Nv_pad * Nv_pad));
    // pixel -> block index, then the transpose stencil
    SRED_TRY(ctx->counts.ensure(sizeof(int) * (size_t)n));
    int *d_pix2v = ctx->counts.as<int>();
    int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_fill_int, dim3(blocks), dim3(256), 0, ctx->stream, d_pix2v, n, -1);
    hipLaunchKernelGGL(k_pix2v, dim3((Nv + 255) / 256), dim3(256), 0, ctx->stream, d_idx, Nv, d_pix2v);
    std::vector<int> dx, dy, dxy(2 * K);
    footprint(g.bw, g.bh, &dx, &dy);
    for (int k = 0; k < K; k++) { dxy[2 * k] = dx[k]; dxy[2 * k + 1] = dy[k]; }
    ctx->host_tmpl.clear();  // the buffer no longer holds a block template
    SRED_TRY(ctx->tmpl.ensure(sizeof(int) * 2 * K));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    SRED_HIP_CHECK(hipMemcpy(ctx->tmpl.ptr, dxy.data(), sizeof(int) * 2 * K, hipMemcpyHostToDevice));
    hipLaunchKernelGGL(k_adjoint_scatter, dim3(blocks), dim3(256), 0, ctx->stream, d_Gc, d_pix2v,
                       ctx->tmpl.as<int>(), K, K16, g.w, g.h, g.brw, g.brh, d_grad);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

}  // namespace sred
