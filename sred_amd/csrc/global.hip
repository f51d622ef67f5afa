// global.hip -- global repetition: all-pairs block comparison as a dense
// contraction with the per-pixel aggregation fused into its epilogue.
//
// Reference: kernelGetGlobal{Pearson,Ssim,Nrmse,AbsDiffStds}2D.cl (one
// work-item per reference pixel looping over every test pixel, gathering
// both blocks per pair) driven by CLUtils.getGlobal*2D (CLUtils.java:
// 1680-2156).  Here the work is re-expressed as:
//
//   1. compaction  -- interior pixels that pass the relevance threshold
//                     (kernel line 33/67: var < threshold skips) in row-major
//                     order, so rows and columns are the same valid set;
//   2. gather      -- each valid block, mean-subtracted and scaled, becomes
//                     one row of U (N x K_pad): FP32 (k-major) or split BF16
//                     hi/lo (fragment-major: MFMA operand order), plus a
//                     float4 of per-block scalars;
//   3. contraction -- U * U^T tile by tile (FP32 VALU or bf16x3 MFMA); each
//                     dot product goes straight into the metric's epilogue
//                     (exp weight, normalisation, truncation, row sums), so
//                     the N x N similarity matrix never leaves registers;
//   4. reduce      -- per-chunk row partials summed in FP64 per pixel;
//   5. finalize    -- CLUtils' host division, inversion, mask, normalise.
//
// Epilogue algebra (exact rewrites of the reference's per-pair formulas):
//   w_ij = exp(-(s_i-s_j)^2 / (h+eps)) = exp2(-(t_i - t_j)^2), t = s*sqrt(log2 e/(h+eps))
//   Pearson: u = (x-mu)/(s*sqrt(K-1)), r = u_i.u_j / (1 + delta),
//            delta = eps/(s_i s_j) (= cov/(s_i s_j + eps)); the term
//            max(r, 0) w = max(u_i.u_j, 0) exp2(-(t_i-t_j)^2 - log2(1+delta));
//            flat blocks have u = 0, so pairs with one flat block give 0 as in
//            the reference; a flat row gets the count of flat columns
//            (both-flat pairs contribute weight 1).
//   SSIM:    u = (x-mu)/sqrt(K-1) (u_i.u_j = cov); flat blocks carry an
//            infinite variance so their term is 0; flat rows as Pearson.
//   NRMSE:   u = (x-mu)/sqrt(K), mse = |u_i|^2 + |u_j|^2 - 2 u_i.u_j.
//   AbsDiff: |s_i - s_j| w_ij = |t_i - t_j| w_ij / sqrt(log2 e/(h+eps)).
// The weight sums Wsum_i = sum_j w_ij depend on the sigmas only and are
// computed apart (weights.hip); the contraction accumulates the similarity
// sums S only.
#include <float.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <type_traits>

#include "sred_internal.h"

namespace sred {

constexpr float kPadT = 1e30f;  // t of padding columns: their weight is exp2(-inf) = 0
// Pearson fast path: with every sigma of a tile >= kFastSigma (or 0),
// delta = eps/(s_i s_j) <= 3.1e-4 and exp(-delta) differs from 1/(1 + delta)
// by delta^2/2 <= 5e-8 relative, so log2(1 + delta) in the exponent becomes
// delta log2(e), one product of a row and a column constant; other tiles use
// the exact form.
constexpr float kFastSigma = 0.018f;
constexpr float kLn2f = 0.693147180559945309f;



// Zero device words with a kernel (a graph can hold it, unlike a memset node).
__global__ void k_zero(unsigned int *p, int64_t nwords)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < nwords; i += (int64_t)gridDim.x * blockDim.x)
        p[i] = 0u;
}

// ---------------------------------------------------------------------------
// 1. compaction (deterministic, row-major order preserved)
// ---------------------------------------------------------------------------
constexpr int kCompactItems = 4;
constexpr int kCompactBlock = 256;
constexpr int kCompactTile = kCompactItems * kCompactBlock;

__device__ __forceinline__ void valid_flags(const float *__restrict__ stds, int64_t p, int64_t n, int w, int h, int brw,
                                            int brh, int d, int brd, double thr_d, int drop_flat, int *valid,
                                            int *flat)
{
#pragma clang fp contract(off)
    *valid = 0;
    *flat = 0;
    if (p >= n) return;
    const int64_t t = p / w;
    const int x = (int)(p - t * w), y = (int)(t % h), z = (int)(t / h);
    if (x < brw || x >= w - brw || y < brh || y >= h - brh || z < brd || z >= d - brd) return;
    float s = stds[p];
    float v = s * s;
    if ((double)v < thr_d) return;  // kernelGetGlobalPearson2D.cl:33,67 (double threshold literal)
    if (drop_flat && s == 0.0f) return;
    *valid = 1;
    *flat = (s == 0.0f);
}

// counts[b] = (valid, flat) pixels of tile b; smax[b] = largest sigma of its valid pixels
__global__ __launch_bounds__(kCompactBlock) void k_compact_count(const float *__restrict__ stds, int w, int h, int brw,
                                                                 int brh, int d, int brd, double thr_d,
                                                                 int drop_flat, int2 *__restrict__ counts,
                                                                 float *__restrict__ smax)
{
    const int64_t n = (int64_t)w * h * d;
    const int64_t base = (int64_t)blockIdx.x * kCompactTile + threadIdx.x * kCompactItems;
    int cv = 0, cf = 0;
    float mx = 0.0f;
    for (int i = 0; i < kCompactItems; i++) {
        int v, f;
        valid_flags(stds, base + i, n, w, h, brw, brh, d, brd, thr_d, drop_flat, &v, &f);
        cv += v;
        cf += f;
        if (v) mx = fmaxf(mx, stds[base + i]);
    }
    __shared__ int sv[kCompactBlock], sf[kCompactBlock];
    __shared__ float sm[kCompactBlock];
    sv[threadIdx.x] = cv;
    sf[threadIdx.x] = cf;
    sm[threadIdx.x] = mx;
    __syncthreads();
    for (int s = kCompactBlock / 2; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            sv[threadIdx.x] += sv[threadIdx.x + s];
            sf[threadIdx.x] += sf[threadIdx.x + s];
            sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        counts[blockIdx.x] = make_int2(sv[0], sf[0]);
        smax[blockIdx.x] = sm[0];
    }
}

// Exclusive scan of block counts (one workgroup); totals at index nblocks.
// offsets[b] = exclusive prefix of counts; offsets[nblocks] = totals;
// offsets[nblocks + 1].x = bits of the largest valid sigma
__global__ __launch_bounds__(1024) void k_compact_scan(const int2 *__restrict__ counts, const float *__restrict__ smax,
                                                       int nblocks, int2 *__restrict__ offsets)
{
    __shared__ int sv[1024], sf[1024];
    __shared__ int carry_v, carry_f;
    if (threadIdx.x == 0) { carry_v = 0; carry_f = 0; }
    __syncthreads();
    for (int b0 = 0; b0 < nblocks; b0 += 1024) {
        int i = b0 + threadIdx.x;
        int2 c = i < nblocks ? counts[i] : make_int2(0, 0);
        sv[threadIdx.x] = c.x;
        sf[threadIdx.x] = c.y;
        __syncthreads();
        for (int s = 1; s < 1024; s <<= 1) {  // Hillis-Steele inclusive scan
            int av = threadIdx.x >= s ? sv[threadIdx.x - s] : 0;
            int af = threadIdx.x >= s ? sf[threadIdx.x - s] : 0;
            __syncthreads();
            sv[threadIdx.x] += av;
            sf[threadIdx.x] += af;
            __syncthreads();
        }
        if (i < nblocks) offsets[i] = make_int2(carry_v + sv[threadIdx.x] - c.x, carry_f + sf[threadIdx.x] - c.y);
        __syncthreads();
        if (threadIdx.x == 1023) { carry_v += sv[1023]; carry_f += sf[1023]; }
        __syncthreads();
    }
    if (threadIdx.x == 0) offsets[nblocks] = make_int2(carry_v, carry_f);
    // largest sigma (a strided pass, then a tree over the 1024 threads)
    float mx = 0.0f;
    for (int i = threadIdx.x; i < nblocks; i += 1024) mx = fmaxf(mx, smax[i]);
    __shared__ float sm[1024];
    sm[threadIdx.x] = mx;
    __syncthreads();
    for (int s = 512; s > 0; s >>= 1) {
        if (threadIdx.x < s) sm[threadIdx.x] = fmaxf(sm[threadIdx.x], sm[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) offsets[nblocks + 1] = make_int2(__float_as_int(sm[0]), 0);
}

__global__ __launch_bounds__(kCompactBlock) void k_compact_scatter(const float *__restrict__ stds, int w, int h,
                                                                   int brw, int brh, int d, int brd, double thr_d,
                                                                   int drop_flat, const int2 *__restrict__ offsets,
                                                                   int *__restrict__ idx)
{
    const int64_t n = (int64_t)w * h * d;
    const int64_t base = (int64_t)blockIdx.x * kCompactTile + threadIdx.x * kCompactItems;
    int flags[kCompactItems], cv = 0;
    for (int i = 0; i < kCompactItems; i++) {
        int f;
        valid_flags(stds, base + i, n, w, h, brw, brh, d, brd, thr_d, drop_flat, &flags[i], &f);
        cv += flags[i];
    }
    __shared__ int sv[kCompactBlock];
    sv[threadIdx.x] = cv;
    __syncthreads();
    for (int s = 1; s < kCompactBlock; s <<= 1) {
        int a = threadIdx.x >= s ? sv[threadIdx.x - s] : 0;
        __syncthreads();
        sv[threadIdx.x] += a;
        __syncthreads();
    }
    int pos = offsets[blockIdx.x].x + sv[threadIdx.x] - cv;
    for (int i = 0; i < kCompactItems; i++)
        if (flags[i]) idx[pos++] = (int)(base + i);
}

// ---------------------------------------------------------------------------
// 2. gather
// ---------------------------------------------------------------------------

// Per-block row scale of u for each metric.
__device__ __forceinline__ float row_scale(int metric, float sigma, float inv_sqrt_km1, float inv_sqrt_k)
{
    if (metric == SRED_PEARSON) return sigma > 0.0f ? inv_sqrt_km1 / sigma : 0.0f;
    if (metric == SRED_SSIM) return inv_sqrt_km1;
    return inv_sqrt_k;  // NRMSE
}

// Split-BF16 blocks in fragment-major order: U[g][ks][lane] (one uint4 = 8
// bf16 each), g = block / 16, lane = 16 * kq + block % 16, holding elements
// k = 32 ks + 8 kq .. 32 ks + 8 kq + 7.  That is exactly the operand fragment
// of lane `lane` of v_mfma_f32_16x16x32_bf16 for k-step ks, for rows (A) and
// columns (B) alike, so every fragment load -- global or LDS -- is 64
// consecutive uint4 per wave, and a 64-column tile is one contiguous 4*KS KiB
// run that can be copied straight into LDS.
//
// Reflection (reflect != 0, blocks of K = 32m + 1 pixels): the block vectors
// u are centred, so they are orthogonal to 1/sqrt(K) (1, ..., 1).  The
// Householder reflection H that swaps that unit vector with e_{K-1} keeps
// every dot product and sends u to H u, whose last component is
// s / sqrt(K) with s = sum_k u_k (zero but for the rounding of the centring,
// so dropping it changes a dot product by s_i s_j / K, second order in that
// rounding).  The other components are u_k + c,
//   c = (u_{K-1} - s / sqrt(K)) / (sqrt(K) - 1),
// one shift for the whole block; the K - 1 of them fit m k-steps of 32.
__device__ __forceinline__ float reflect_shift(const float *img, int p, const int *offs, int K, float mu, float sc)
{
    float s = 0.0f;
    for (int k = 0; k < K; k++) s += (img[p + offs[k]] - mu) * sc;
    const float ulast = (img[p + offs[K - 1]] - mu) * sc;
    const float rk = sqrtf((float)K);
    return (ulast - s / rk) / (rk - 1.0f);
}

__global__ __launch_bounds__(256) void k_gather_bf16(int metric, const float *__restrict__ img,
                                                     const float *__restrict__ means, const float *__restrict__ stds,
                                                     const int *__restrict__ idx, int Nv, int Nv_pad, int K, int KS,
                                                     const int *__restrict__ offs, float inv_sqrt_km1,
                                                     float inv_sqrt_k, int reflect, uint4 *__restrict__ Uhi,
                                                     uint4 *__restrict__ Ulo)
{
    const int per_group = KS * 64;  // uint4 per 16 blocks
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Nv_pad * KS * 4) return;
    const int g = (int)(t / per_group), rem = (int)(t - (int64_t)g * per_group);
    const int lane = rem & 63;
    const int v = g * 16 + (lane & 15), q = (rem >> 6) * 4 + (lane >> 4);
    unsigned short hi[8], lo[8];
    if (v < Nv) {
        int p = idx[v];
        float mu = means[p];
        float sc = row_scale(metric, stds[p], inv_sqrt_km1, inv_sqrt_k);
        const int Kd = reflect ? K - 1 : K;  // dimensions stored
        const float c = reflect ? reflect_shift(img, p, offs, K, mu, sc) : 0.0f;
        for (int e = 0; e < 8; e++) {
            int k = q * 8 + e;
            float u = 0.0f;
            if (k < Kd) u = (img[p + offs[k]] - mu) * sc + c;  // test_block[i] - test_mean (kernel line 84-87)
            hi[e] = f2bf(u);
            lo[e] = f2bf(u - bf2f(hi[e]));
        }
    } else {
        for (int e = 0; e < 8; e++) hi[e] = lo[e] = 0;
    }
    uint4 H, L;
    H.x = hi[0] | ((unsigned)hi[1] << 16); H.y = hi[2] | ((unsigned)hi[3] << 16);
    H.z = hi[4] | ((unsigned)hi[5] << 16); H.w = hi[6] | ((unsigned)hi[7] << 16);
    L.x = lo[0] | ((unsigned)lo[1] << 16); L.y = lo[2] | ((unsigned)lo[3] << 16);
    L.z = lo[4] | ((unsigned)lo[5] << 16); L.w = lo[6] | ((unsigned)lo[7] << 16);
    Uhi[t] = H;
    Ulo[t] = L;
}

// FP32 rows, k-major: UT[k][v] for k < K4 (K rounded up to 4).
__global__ __launch_bounds__(256) void k_gather_f32(int metric, const float *__restrict__ img,
                                                    const float *__restrict__ means, const float *__restrict__ stds,
                                                    const int *__restrict__ idx, int Nv, int Nv_pad, int K, int K4,
                                                    const int *__restrict__ offs, float inv_sqrt_km1,
                                                    float inv_sqrt_k, float *__restrict__ UT)
{
    int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)Nv_pad * K4) return;
    int k = (int)(t / Nv_pad), v = (int)(t - (int64_t)k * Nv_pad);
    float u = 0.0f;
    if (v < Nv && k < K) {
        int p = idx[v];
        u = (img[p + offs[k]] - means[p]) * row_scale(metric, stds[p], inv_sqrt_km1, inv_sqrt_k);
    }
    UT[t] = u;
}

// Per-block scalars (see the layout of the epilogue constants below).
__global__ __launch_bounds__(256) void k_aux(int metric, const float *__restrict__ img, const float *__restrict__ means,
                                             const float *__restrict__ stds, const int *__restrict__ idx, int Nv,
                                             int Nv_pad, int K, const int *__restrict__ offs, float tscale,
                                             float inv_sqrt_km1, float inv_sqrt_k, int split_bf16, int reflect,
                                             float4 *__restrict__ aux)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Nv_pad) return;
    float4 a = make_float4(kPadT, 0.0f, 0.0f, 0.0f);
    if (v >= Nv) {
        if (metric == SRED_SSIM) a.w = INFINITY;
        if (metric == SRED_NRMSE) a.y = -INFINITY;
        aux[v] = a;
        return;
    }
    int p = idx[v];
    float s = stds[p], mu = means[p];
    a.x = s * tscale;
    if (metric == SRED_PEARSON) {
        a.y = s > 0.0f ? 1.0f / s : 0.0f;
        a.z = (s > 0.0f && s < kFastSigma) ? 1.0f : 0.0f;
    } else if (metric == SRED_SSIM) {
        a.y = mu;
        a.z = mu * mu;
        a.w = s == 0.0f ? INFINITY : s * s;
    } else if (metric == SRED_NRMSE) {
        // |u|^2 as the contraction sees it: with split BF16 the products are
        // hi*hi + hi*lo + lo*hi, so the norm omits lo*lo too and identical
        // blocks give mse = 0 up to FP32 rounding
        float sc = row_scale(metric, s, inv_sqrt_km1, inv_sqrt_k), n = 0.0f;
        const int Kd = reflect ? K - 1 : K;  // the reflected blocks' components (k_gather_bf16)
        const float c = reflect ? reflect_shift(img, p, offs, K, mu, sc) : 0.0f;
        for (int k = 0; k < Kd; k++) {
            float u = (img[p + offs[k]] - mu) * sc + c;
            if (split_bf16) {
                float hi = bf2f(f2bf(u)), lo = bf2f(f2bf(u - hi));
                n += hi * hi + 2.0f * hi * lo;
            } else {
                n += u * u;
            }
        }
        a.y = n;
    }
    aux[v] = a;
}

// tile_small[g] = 1 if any block of rows/columns [32g, 32g+32) has 0 < sigma < kFastSigma
__global__ __launch_bounds__(256) void k_tile_small(const float4 *__restrict__ aux, int ngroups,
                                                    int *__restrict__ tile_small)
{
    int g = (blockIdx.x * blockDim.x + threadIdx.x) / 32;
    int lane32 = threadIdx.x & 31;
    if (g >= ngroups) return;
    int small = aux[g * 32 + lane32].z != 0.0f;
    unsigned long long b = __ballot(small);
    unsigned int half = (threadIdx.x & 32) ? (unsigned int)(b >> 32) : (unsigned int)b;
    if (lane32 == 0) tile_small[g] = half != 0u;
}

// ---------------------------------------------------------------------------
// 3. epilogue
// ---------------------------------------------------------------------------
// Row constants R and column constants C from the per-block scalars.  The
// per-block float4 is already the column form C (so a column tile's constants
// can be copied into LDS as they are):
//   Pearson: R = (t, eps log2(e)/s),  C = (t, 1/s)
//   SSIM:    R = (t, 2mu, mu^2, s^2+c2), C = (t, mu, mu^2, s^2)   (inf for flat)
//   NRMSE:   R = (t, |u|^2),          C = (t, |u|^2)              (-inf for padding; also
//            the weighted NRMSE, whose padding terms have weight 0)
//   AbsDiff: R = (t),                 C = (t)
template <int M> __device__ __forceinline__ float4 row_const(float4 a, const EpiConst &e)
{
    if (M == SRED_PEARSON) return make_float4(a.x, e.pk * a.y, 0.0f, 0.0f);
    if (M == SRED_SSIM) return make_float4(a.x, 2.0f * a.y, a.z, a.w + e.c2);
    return a;
}
template <int M> __device__ __forceinline__ float4 col_const(float4 a, const EpiConst &) { return a; }

// FOLD (the symmetric kernel): the MFMA accumulators did not start at zero,
// so `dot` already holds the metric's affine part (see sym_row_setup):
//   SSIM:  dot = u_i.u_j + c2/2, and R.y = 4 m_i (the factor 2 of 2cov + c2
//          moved into the luminance term: (4 m_i m_j + 2 c1)(cov + c2/2));
//   NRMSE: dot = |u_i|^2 - 2 u_i.u_j (row fragments scaled by -2).
template <int M, bool EXACT = true, bool COLS = false, bool FOLD = false>
__device__ __forceinline__ void epilogue(float dot, const float4 &R, const float4 &C, float &S, const EpiConst &e,
                                         float *cS = nullptr)
{
    const float d = R.x - C.x;
    if (M == SRED_PEARSON) {
        const float p = R.y * C.y;  // delta log2(e)
        // exact: log2(1 + delta); fast: delta log2(e)
        const float x = EXACT ? fmaf(d, d, __builtin_amdgcn_logf(fmaf(p, kLn2f, 1.0f))) : fmaf(d, d, p);
        const float wt = __builtin_amdgcn_exp2f(-x);
        // max(dot, 0) as one v_med3 (fmaxf would first canonicalise the MFMA result)
        const float r = __builtin_amdgcn_fmed3f(dot, 0.0f, FLT_MAX);
        S = fmaf(r, wt, S);
        if (COLS) *cS = fmaf(r, wt, *cS);
    } else if (M == SRED_SSIM) {
        const float w = __builtin_amdgcn_exp2f(-(d * d));
        float a = fmaf(R.y, C.y, FOLD ? 2.0f * e.c1 : e.c1);
        float b = FOLD ? dot : fmaf(2.0f, dot, e.c2);
        float c = fmaf(R.z, C.z, e.c1);
        float t = fmaxf((a * b) * __builtin_amdgcn_rcpf(c * (R.w + C.w)) * w, 0.0f);
        S += t;
        if (COLS) *cS += t;
    } else if (M == SRED_NRMSE) {  // unweighted (kernelGetGlobalNrmse2D.cl)
        float mse = FOLD ? dot + C.y : fmaf(-2.0f, dot, R.y + C.y);
        float t = __builtin_amdgcn_sqrtf(fmaxf(mse, 0.0f));
        S += t;
        if (COLS) *cS += t;
    } else if (M == kNrmseWeighted) {  // sqrt(mse) * weight (kernelGetGlobalNrmse3D.cl:111)
        const float w = __builtin_amdgcn_exp2f(-(d * d));
        float mse = FOLD ? dot + C.y : fmaf(-2.0f, dot, R.y + C.y);
        float t = __builtin_amdgcn_sqrtf(fmaxf(mse, 0.0f)) * w;
        S += t;
        if (COLS) *cS += t;
    } else {  // ABSDIFF
        const float w = __builtin_amdgcn_exp2f(-(d * d));
        S = fmaf(fabsf(d), w, S);
        if (COLS) *cS = fmaf(fabsf(d), w, *cS);
    }
}

// ---------------------------------------------------------------------------
// 3a. split-BF16 MFMA contraction (v_mfma_f32_16x16x32_bf16).
// Workgroup: 4 waves; each wave owns RG*16 rows whose A fragments (hi and lo)
// stay in registers for the whole column sweep.  Column tiles of 64 blocks
// are staged in LDS (hi and lo) and shared by the 4 waves.  MFMA lane layout
// (16x16 output): lane l holds column l%16 and rows 4*(l/16)+0..3; operand
// lanes hold row/column l%16 and k-group l/16 (8 consecutive k).  The k
// assignment is identical for both operands, so it cancels in the dot.
// Workgroup ids are chunk-major: concurrently resident workgroups sweep the
// same column chunk (sized to stay in each XCD's L2).
// ---------------------------------------------------------------------------
constexpr int kBN = 64;
// rows per wave = 16*RG: 4 row groups while the A fragments fit (K_pad <= 128),
// 2 beyond; every B fragment read from LDS then serves 64 (or 32) rows.
// (SSIM carries four row constants per row, so it keeps 4 row groups only up to K_pad = 96).
constexpr int rg_for(int metric, int ks) { return ks <= (metric == SRED_SSIM ? 3 : 4) ? 4 : 2; }
// the symmetric kernel also carries column sums and a second accumulator set
// (software pipelining): 2 row groups up to K_pad = 256.  Beyond K_pad = 160
// the registers then spill a little, but every B fragment read from LDS feeds
// twice the MFMAs; measured on a 768^2 lattice image, against one row group:
// 15x17 blocks (K_pad 192) Pearson 147 -> 113 ms, 17x17 (224) 166 -> 134 ms,
// 19x19 (256) 319 -> 305 ms; SSIM and NRMSE gain 15-30%.  K_pad 288-384
// (21x21, 23x23 blocks) keeps one row group: their tiles leave one workgroup
// per CU.  Four row groups (half the tile copies, folds and barriers per
// pair) where that was measured faster at 1024^2: Pearson at K_pad = 32
// (5x5: 157.5 -> 145.5 ms) and NRMSE at K_pad = 64 (9x9: 189 -> 172 ms); the
// other cases spill more and ran slower with four (Pearson 9x9 +5%, SSIM
// 9x9 +47%, NRMSE 5x5 +2%, 11x11 +10%, 13x13 2.2x).
constexpr int rg_sym(int metric, int ks)
{
    return ((metric == SRED_PEARSON && ks == 1) || (metric == SRED_NRMSE && ks == 2)) ? 4 : ks <= 8 ? 2 : 1;
}

template <int M, int KS>
__global__ __launch_bounds__(256, 2) void k_contract_mfma(const uint4 *__restrict__ Uhi, const uint4 *__restrict__ Ulo,
                                                       const float4 *__restrict__ aux,
                                                       const int *__restrict__ tile_small, int nrowtiles,
                                                       int cols_per_chunk, int ncols, float *__restrict__ partial,
                                                       int Nv_pad, EpiConst ec)
{
    constexpr int RG = rg_for(M, KS);
    constexpr int BM = 64 * RG;
    constexpr int TILE = 4 * KS * 64;  // uint4 of one 64-column tile (fragment-major)
    __shared__ uint4 lds_hi[TILE];
    __shared__ uint4 lds_lo[TILE];

    const int rt = blockIdx.x % nrowtiles, ch = blockIdx.x / nrowtiles;
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int li = lane & 15, lk = lane >> 4;
    const int rowbase = rt * BM + wave * 16 * RG;

    bf16x8 ahi[RG][KS], alo[RG][KS];
#pragma unroll
    for (int rg = 0; rg < RG; rg++)
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            size_t o = ((size_t)(rowbase / 16 + rg) * KS + ks) * 64 + lane;
            uint4 h = Uhi[o], l = Ulo[o];
            ahi[rg][ks] = __builtin_bit_cast(bf16x8, h);
            alo[rg][ks] = __builtin_bit_cast(bf16x8, l);
        }
    float4 R[RG][4];
#pragma unroll
    for (int rg = 0; rg < RG; rg++)
#pragma unroll
        for (int r = 0; r < 4; r++) R[rg][r] = row_const<M>(aux[rowbase + rg * 16 + lk * 4 + r], ec);
    // rows of this wave with a small sigma force the exact Pearson epilogue
    bool row_small = false;
    if (M == SRED_PEARSON)
        for (int g = 0; g < RG / 2; g++) row_small |= tile_small[rowbase / 32 + g] != 0;

    // per-lane FP32 row sums over the chunk (at most cols_per_chunk/16 terms each)
    float accS[RG][4];
#pragma unroll
    for (int rg = 0; rg < RG; rg++)
#pragma unroll
        for (int r = 0; r < 4; r++) accS[rg][r] = 0.0f;

    const int c_begin = ch * cols_per_chunk;
    const int c_end = min(c_begin + cols_per_chunk, ncols);
    for (int c0 = c_begin; c0 < c_end; c0 += kBN) {
        __syncthreads();
        const size_t t0 = (size_t)(c0 / 16) * KS * 64;
        for (int i = tid; i < TILE; i += 256) {
            lds_hi[i] = Uhi[t0 + i];
            lds_lo[i] = Ulo[t0 + i];
        }
        __syncthreads();
        const bool exact = row_small || (M == SRED_PEARSON && (tile_small[c0 / 32] | tile_small[c0 / 32 + 1]) != 0);
#pragma unroll 1
        for (int cg = 0; cg < 4; cg++) {
            const int col = cg * 16 + li;
            const float4 C = col_const<M>(aux[c0 + col], ec);
            floatx4 acc[RG];
#pragma unroll
            for (int rg = 0; rg < RG; rg++) acc[rg] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                bf16x8 bh = __builtin_bit_cast(bf16x8, lds_hi[(cg * KS + ks) * 64 + lane]);
                bf16x8 bl = __builtin_bit_cast(bf16x8, lds_lo[(cg * KS + ks) * 64 + lane]);
#pragma unroll
                for (int rg = 0; rg < RG; rg++) {
                    acc[rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[rg][ks], bh, acc[rg], 0, 0, 0);
                    acc[rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[rg][ks], bl, acc[rg], 0, 0, 0);
                    acc[rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[rg][ks], bh, acc[rg], 0, 0, 0);
                }
            }
            if (exact) {
#pragma unroll
                for (int rg = 0; rg < RG; rg++)
#pragma unroll
                    for (int r = 0; r < 4; r++) epilogue<M, true>(acc[rg][r], R[rg][r], C, accS[rg][r], ec);
            } else {
#pragma unroll
                for (int rg = 0; rg < RG; rg++)
#pragma unroll
                    for (int r = 0; r < 4; r++) epilogue<M, false>(acc[rg][r], R[rg][r], C, accS[rg][r], ec);
            }
        }
    }
#pragma unroll
    for (int rg = 0; rg < RG; rg++)
#pragma unroll
        for (int r = 0; r < 4; r++) {
            float s = accS[rg][r];
#pragma unroll
            for (int m = 1; m < 16; m <<= 1) s += __shfl_xor(s, m);
            if (li == 0) partial[(size_t)ch * Nv_pad + rowbase + rg * 16 + lk * 4 + r] = s;
        }
}

// Folds each metric's affine part into the symmetric kernel's MFMAs (one
// vector instruction less per pair in the epilogue<..., FOLD = true>):
//   SSIM:  accumulators start at c2/2, R.y = 2 m_i becomes 4 m_i;
//   NRMSE (and the weighted 3D form): row fragments times -2 (a power of
//          two: exact), accumulators start at the row's |u_i|^2;
//   Pearson: accumulators start at 0.
// Measured against the unfolded form: 1024^2 11x11 SSIM 274.6 -> 268.7 ms,
// NRMSE 256.2 -> 251.6 ms, 768^2 13x13 -2% / -1%.  SSIM beyond K_pad = 128
// keeps the unfolded form (the start values cost registers there: 17x17
// spilled more and ran 2% slower).
template <int M, int KS> constexpr bool sym_fold() { return M == SRED_SSIM ? KS <= 4 : M != SRED_PEARSON; }
template <int M, int KS, int RG>
__device__ __forceinline__ void sym_row_setup(bf16x8 (&ahi)[RG][KS], bf16x8 (&alo)[RG][KS], float4 (&R)[RG][4],
                                              floatx4 (&cinit)[RG], const EpiConst &ec)
{
#pragma unroll
    for (int rg = 0; rg < RG; rg++) {
        if (M == SRED_NRMSE || M == kNrmseWeighted) {
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                ahi[rg][ks] = bf16x8_scaled(ahi[rg][ks], -2.0f);
                alo[rg][ks] = bf16x8_scaled(alo[rg][ks], -2.0f);
            }
            cinit[rg] = floatx4{R[rg][0].y, R[rg][1].y, R[rg][2].y, R[rg][3].y};
        } else if (M == SRED_SSIM && sym_fold<M, KS>()) {
            const float h = 0.5f * ec.c2;
            cinit[rg] = floatx4{h, h, h, h};
#pragma unroll
            for (int r = 0; r < 4; r++) R[rg][r].y *= 2.0f;
        } else {
            cinit[rg] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        }
    }
}

// ---------------------------------------------------------------------------
// One 64-column tile of one wave (16*RG rows): the MFMAs of column group
// cg+1 are issued before the epilogue of group cg, so the matrix core works
// on the next group while the vector ALU finishes the current one; the
// scheduling hints interleave one MFMA with a few VALU instructions.  COLS
// adds the column sums (off-diagonal panel pairs): per lane over its rows for
// each of the 4 column groups, stored as one float4 per lane at the end of
// the tile (the k-groups and waves are summed when the tile is folded, so no
// cross-lane operation is needed).  B fragments are read from the
// fragment-major LDS tile: 64 consecutive uint4 per (group, k-step).
// Pearson and the unweighted NRMSE with K <= 96 read the B fragments of
// group cg+2 into a second register set while group cg+1's MFMAs and group
// cg's epilogue run, so no MFMA waits on LDS latency (measured 2.8% faster at
// 1024^2, 11x11); the other cases read each group's fragments just before
// its MFMAs (with the second set SSIM spills and runs 4% slower, and so do
// all metrics beyond K = 96).
template <int M, int KS, int RG, bool EXACT, bool COLS>
__device__ __forceinline__ void sym_tile(const uint4 *lds_hi, const uint4 *lds_lo, const float4 *ccon,
                                         const bf16x8 (&ahi)[RG][KS], const bf16x8 (&alo)[RG][KS],
                                         const float4 (&R)[RG][4], const floatx4 (&cinit)[RG], float (&accS)[RG][4],
                                         float4 *stage_w, int lane, const EpiConst &ec)
{
    constexpr bool PF = (M == SRED_PEARSON || M == SRED_NRMSE) && KS <= 3;
    const int li = lane & 15;
    floatx4 acc[2][RG];
    float colsum[4];
    uint4 bh[PF ? 2 : 1][KS], bl[PF ? 2 : 1][KS];
    if (PF) {
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            bh[0][ks] = lds_hi[ks * 64 + lane];
            bl[0][ks] = lds_lo[ks * 64 + lane];
        }
    }
#pragma unroll
    for (int cg = -1; cg < 4; cg++) {
        if (PF && cg + 2 < 4) {  // B fragments of group cg+2, into the set group cg used
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                bh[(cg + 2) & 1][ks] = lds_hi[((cg + 2) * KS + ks) * 64 + lane];
                bl[(cg + 2) & 1][ks] = lds_lo[((cg + 2) * KS + ks) * 64 + lane];
            }
        }
        if (cg + 1 < 4) {  // issue group cg+1
            floatx4(&a)[RG] = acc[(cg + 1) & 1];
#pragma unroll
            for (int rg = 0; rg < RG; rg++) a[rg] = cinit[rg];
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                const bf16x8 h = __builtin_bit_cast(
                    bf16x8, PF ? bh[(cg + 1) & 1][ks] : lds_hi[((cg + 1) * KS + ks) * 64 + lane]);
                const bf16x8 l = __builtin_bit_cast(
                    bf16x8, PF ? bl[(cg + 1) & 1][ks] : lds_lo[((cg + 1) * KS + ks) * 64 + lane]);
#pragma unroll
                for (int rg = 0; rg < RG; rg++) {
                    a[rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[rg][ks], h, a[rg], 0, 0, 0);
                    a[rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ahi[rg][ks], l, a[rg], 0, 0, 0);
                    a[rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(alo[rg][ks], h, a[rg], 0, 0, 0);
                }
            }
        }
        if (cg >= 0) {  // epilogue of group cg
            const float4 C = ccon[cg * 16 + li];
            float cs = 0.0f;
#pragma unroll
            for (int rg = 0; rg < RG; rg++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    epilogue<M, EXACT, COLS, sym_fold<M, KS>()>(acc[cg & 1][rg][r], R[rg][r], C, accS[rg][r], ec,
                                                                &cs);
            colsum[cg] = cs;
        }
        if (M != SRED_SSIM && cg >= 0 && cg + 1 < 4) {
            // the prefetch reads first, then per MFMA of the next group two
            // VALU/transcendental ops of this one (measured best for Pearson;
            // SSIM's longer epilogue is left to the compiler's own order,
            // also measured faster)
            if (PF && cg + 2 < 4) __builtin_amdgcn_sched_group_barrier(0x100, 2 * KS, 0);  // DS reads
#pragma unroll
            for (int i = 0; i < 3 * KS * RG; i++) {
                __builtin_amdgcn_sched_group_barrier(0x008, 1, 0);  // MFMA
                __builtin_amdgcn_sched_group_barrier(0x402, 2, 0);  // VALU | TRANS
            }
        }
    }
    if (COLS) stage_w[lane] = make_float4(colsum[0], colsum[1], colsum[2], colsum[3]);
}

// Asynchronous global -> LDS copy of one 64-column tile: 4*KS runs of 64
// uint4 for hi and for lo (contiguous in the fragment-major layout), plus the
// tile's 64 column constants.  Wave w copies runs w, w+4, ...; with
// global_load_lds_dwordx4 the LDS destination is wave-uniform and lane l
// lands 16*l bytes after it, so LDS keeps the global order.  Completion is
// tracked by the issuing wave's vmcnt.
template <int KS>
__device__ __forceinline__ void tile_to_lds(const uint4 *Uhi, const uint4 *Ulo, const float4 *aux, int c0, uint4 *dhi,
                                            uint4 *dlo, float4 *dcon, int wave, int lane)
{
    typedef __attribute__((address_space(3))) void *lds_ptr;
    const size_t t0 = (size_t)(c0 / 16) * KS * 64;
#pragma unroll
    for (int i = 0; i < KS; i++) {
        const int run = wave + 4 * i;
        __builtin_amdgcn_global_load_lds((const void *)(Uhi + t0 + run * 64 + lane), (lds_ptr)(dhi + run * 64), 16, 0,
                                         0);
        __builtin_amdgcn_global_load_lds((const void *)(Ulo + t0 + run * 64 + lane), (lds_ptr)(dlo + run * 64), 16, 0,
                                         0);
    }
    if (wave == 0) __builtin_amdgcn_global_load_lds((const void *)(aux + c0 + lane), (lds_ptr)dcon, 16, 0, 0);
}

// s_waitcnt vmcnt(0) expcnt(0) lgkmcnt(0): this wave's global->LDS copies
// (vmcnt) and its own LDS stores (lgkmcnt) are complete.  The LDS count is
// waited for explicitly too: an explicit wait for the vector-memory count
// alone right before a barrier has been seen to take the place of the
// LDS-count wait the barrier needs, leaving a tile's column-sum stores in
// flight while other waves read them.
__device__ __forceinline__ void wait_vm_all() { __builtin_amdgcn_s_waitcnt(0); }

// Weighted metrics (Pearson, SSIM) skip pairs whose t values are more than
// skip_gap<M>() apart.  Every skipped term is then below 2^-40 in magnitude,
// so a row sum over N blocks changes by less than N * 2^-40 (< 1e-6 for
// N = 10^6, against a self-pair term of about 1):
//   Pearson: max(0, r) <= 1, weight exp2(-gap^2) < 2^-40 with gap^2 = 40;
//   SSIM:    |2cov + c2| <= s_i^2 + s_j^2 + c2 (Cauchy-Schwarz), but the
//            luminance factor (2 m_i m_j + c1) / (m_i^2 m_j^2 + c1) of the
//            product denominator (kernelGetGlobalSsim2D.cl:113) reaches
//            100.5 (at m_i m_j = 0.01, for any real means), so the term is at
//            most 100.5 w and gap^2 = 47 gives 100.5 * 2^-47 < 2^-40.
// The 3D weighted NRMSE (sqrt(mse) * w, unbounded for inputs outside [0, 1])
// and the unweighted NRMSE are never skipped.
template <int M> constexpr float skip_gap() { return M == SRED_SSIM ? 6.8556546f : 6.3245553f; }  // sqrt(47), sqrt(40)
template <int M> constexpr bool skips() { return M == SRED_PEARSON || M == SRED_SSIM; }

// two 64-column tiles in flight (copy of the next overlaps the current one)
// while their LDS fits twice beside the column accumulator at 2 workgroups/CU:
// up to K_pad = 128, where the tiles take 64 KiB and the panel is held to
// 1024 columns (4 KiB accumulator; measured 3% faster than one buffer at
// P = 2048, while two buffers at P = 2048 leave one workgroup per CU and
// run 35% slower)
constexpr bool sym_double_buffer(int ks) { return ks <= 4; }

// ---------------------------------------------------------------------------
// 3a'. symmetric split-BF16 MFMA contraction.  Every per-pair term is
// symmetric in (i, j), so only panel pairs (p, c) with c >= p are evaluated.
// Panels are P rows = P columns of the valid set.  Workgroup (p, c), c > p:
// the P x P rectangle's row sums go to partial[c][rows of p] and its column
// sums to partial[p][columns of c]; workgroup (p, p) evaluates its whole
// diagonal square and writes row sums only.  Every partial slot therefore has
// exactly one writer (deterministic, no atomics), and the FP64 reduction over
// all slots gives the full row sums.  Column sums: per lane over its 16*RG
// rows, across the 4 k-groups by shuffles, across the 4 waves through an LDS
// staging buffer folded into an LDS accumulator for the panel's P columns.
//
// The kernel is persistent: workgroup g takes work items g, g + gridDim, ...
// of a list that may span several problems (SymProblem, one per image of a
// batch); an item is one panel pair of one problem.  Items of one problem are
// contiguous and chunk-major, so the workgroups resident together still share
// a column panel.
//
// Valid blocks are ordered by ascending sigma (so ascending t): for Pearson
// and SSIM a panel pair whose t ranges are more than skip_gap apart writes
// zero partial sums without being evaluated, and inside an evaluated panel
// pair a wave skips each column tile whose t range is that far from its
// rows' range.  next[1] counts the ordered pairs skipped.
//
// The workgroup walks its rectangle as P/BM sweeps of P/64 column tiles
// (step T = tile T % ntiles of sweep T / ntiles).  With double buffering the
// copy of tile T+1 is issued before tile T is computed, into the other of two
// LDS buffers (distinct arrays, so the compiler sees that the reads of one do
// not depend on the copy into the other); one barrier per step then both
// publishes tile T and retires every read of the buffer the next copy reuses.
// ---------------------------------------------------------------------------
template <int M, int KS, bool RECT>
__global__ __launch_bounds__(256, KS <= 8 ? 2 : 1) void k_contract_mfma_sym(const SymProblem *__restrict__ probs, int nprob,
                                                              long long nitems, unsigned long long *__restrict__ next,
                                                              int P, EpiConst ec)
{
    constexpr int RG = rg_sym(M, KS);
    constexpr int BM = 64 * RG;
    constexpr int TILE = 4 * KS * 64;
    constexpr bool DB = sym_double_buffer(KS);
    __shared__ uint4 hi0[TILE], lo0[TILE];
    __shared__ float4 con0[kBN];
    __shared__ uint4 hi1[DB ? TILE : 1], lo1[DB ? TILE : 1];
    __shared__ float4 con1[DB ? kBN : 1];
    __shared__ float4 stage[2][4][64];  // [parity][wave][lane]: column-group sums
    extern __shared__ float colacc[];  // P entries

    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    __shared__ long long s_item;
  for (;;) {
    // next item from the launch's counter, in order (as the hardware would
    // dispatch one workgroup per item); every thread has read the previous
    // s_item and every LDS access of the previous item is done
    wait_vm_all();
    __syncthreads();
    if (tid == 0) s_item = (long long)atomicAdd(next, 1ull);
    __syncthreads();
    const long long item = s_item;
    if (item >= nitems) break;
    // the problem holding this item: the last one whose first item is <= item
    int qlo = 0, qhi = nprob - 1;
    while (qlo < qhi) {
        const int mid = (qlo + qhi + 1) >> 1;
        if (probs[mid].item0 <= item) qlo = mid;
        else qhi = mid - 1;
    }
    const uint4 *__restrict__ Uhi = probs[qlo].uhi;
    const uint4 *__restrict__ Ulo = probs[qlo].ulo;
    const float4 *__restrict__ aux = probs[qlo].aux;
    const int *__restrict__ tile_small = probs[qlo].tsm;
    float *__restrict__ partial = probs[qlo].partial;
    const int Nv_pad = probs[qlo].Nv_pad;
    // RECT: the launch may hold rectangular problems (super-panel pairs); the
    // other instantiation keeps those fields out of its registers (measured
    // 1% at 1024^2, 11x11)
    const int npr = RECT ? probs[qlo].npr : 0, coff = RECT ? probs[qlo].coff : 0;
    float *__restrict__ pcol = (RECT && npr > 0) ? probs[qlo].pcol : partial;
    const int ldc = (RECT && npr > 0) ? probs[qlo].ldc : Nv_pad;
    // (p, c) from the chunk-major index q: triangular q = c(c+1)/2 + p, or
    // rectangular q = c npr + p
    const int q = (int)(item - probs[qlo].item0);
    int c, p;
    if (npr > 0) {
        c = q / npr;
        p = q - c * npr;
    } else {
        c = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
        while ((c + 1) * (c + 2) / 2 <= q) c++;
        while (c * (c + 1) / 2 > q) c--;
        p = q - c * (c + 1) / 2;
    }
    const bool diag = npr == 0 && p == c;
    const int col0 = coff + c * P, row0 = p * P;
    if (skips<M>() && !diag && aux[col0].x - aux[row0 + P - 1].x > skip_gap<M>()) {
        // every pair of the panel pair has a vanishing weight
        for (int i = tid; i < P; i += 256) {
            partial[(size_t)c * Nv_pad + row0 + i] = 0.0f;
            pcol[(size_t)p * ldc + (col0 - coff) + i] = 0.0f;
        }
        if (tid == 0) atomicAdd(next + 1, 2ull * (unsigned long long)P * (unsigned long long)P);
        continue;
    }

    if (!diag)
        for (int i = tid; i < P; i += 256) colacc[i] = 0.0f;

    const int ntiles = P / kBN;  // multiple of 4 (the step count below is even), at most 64
    const int NT = ntiles * (P / BM);
    // per column tile of the panel, held by lane (tile index): its t range
    // and whether it holds a small-sigma block (exact Pearson epilogue); read
    // back with readlane, so the tile loop waits on no memory for them
    float tile_tlo = 0.0f, tile_thi = 0.0f;
    unsigned long long tile_exact = 0ull;
    if (lane < ntiles) {
        const int cc = col0 + lane * kBN;
        if (skips<M>()) {
            tile_tlo = aux[cc].x;
            tile_thi = aux[cc + kBN - 1].x;
        }
        if (M == SRED_PEARSON) tile_exact = (tile_small[cc / 32] | tile_small[cc / 32 + 1]) != 0;
    }
    if (M == SRED_PEARSON) tile_exact = __ballot(tile_exact != 0ull);

    bf16x8 ahi[RG][KS], alo[RG][KS];
    float4 R[RG][4];
    floatx4 cinit[RG];  // MFMA accumulator start values (the folded affine part of the metric)
    float accS[RG][4];
    bool row_small = false;
    int rowbase = 0;
    float wt_lo = 0.0f, wt_hi = 0.0f;  // t range of this wave's rows
    unsigned long long nskip = 0;      // ordered pairs skipped by this wave

    auto fold = [&](int T) {
        // column sums of step T's tile into colacc (after a barrier), spread
        // over all threads: wave g takes column group g; thread (c, kq) sums
        // column 16g+c of k-group kq over the 4 waves, then the 4 k-groups
        // (adjacent lanes) are added by DPP quad permutes
        const int c = (tid >> 2) & 15, kq = tid & 3;
        const float *st = reinterpret_cast<const float *>(&stage[T & 1][0][kq * 16 + c]) + wave;
        float v = ((st[0] + st[256]) + st[512]) + st[768];  // waves 0..3 (64 float4 apart)
        v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0xB1, 0xF, 0xF, true));  // kq ^ 1
        v += __int_as_float(__builtin_amdgcn_mov_dpp(__float_as_int(v), 0x4E, 0xF, 0xF, true));  // kq ^ 2
        if (kq == 0) colacc[(T % ntiles) * kBN + wave * 16 + c] += v;
    };
    auto begin_sweep = [&](int T) {  // A fragments and row constants of sweep T / ntiles
        rowbase = row0 + (T / ntiles) * BM + wave * 16 * RG;
#pragma unroll
        for (int rg = 0; rg < RG; rg++)
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                const size_t o = ((size_t)(rowbase / 16 + rg) * KS + ks) * 64 + lane;
                ahi[rg][ks] = __builtin_bit_cast(bf16x8, Uhi[o]);
                alo[rg][ks] = __builtin_bit_cast(bf16x8, Ulo[o]);
            }
        const int lk = lane >> 4;
#pragma unroll
        for (int rg = 0; rg < RG; rg++)
#pragma unroll
            for (int r = 0; r < 4; r++) R[rg][r] = row_const<M>(aux[rowbase + rg * 16 + lk * 4 + r], ec);
        sym_row_setup<M, KS, RG>(ahi, alo, R, cinit, ec);
        row_small = false;
        if (M == SRED_PEARSON)
            for (int g = 0; g < RG / 2; g++) row_small |= tile_small[rowbase / 32 + g] != 0;
        wt_lo = aux[rowbase].x;
        wt_hi = aux[rowbase + 16 * RG - 1].x;
#pragma unroll
        for (int rg = 0; rg < RG; rg++)
#pragma unroll
            for (int r = 0; r < 4; r++) accS[rg][r] = 0.0f;
    };
    auto end_sweep = [&]() {  // row sums of the sweep -> partial[c][rows]
        const int li = lane & 15, lk = lane >> 4;
#pragma unroll
        for (int rg = 0; rg < RG; rg++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                float s = accS[rg][r];
#pragma unroll
                for (int m = 1; m < 16; m <<= 1) s += __shfl_xor(s, m);
                if (li == 0) partial[(size_t)c * Nv_pad + rowbase + rg * 16 + lk * 4 + r] = s;
            }
    };
    // SKIP (a std::bool_constant): whether this panel pair may hold tiles
    // whose weights vanish (decided once per panel pair below), so pairs
    // that cannot skip run a tile loop without the test
    auto compute = [&](auto skip_tag, int T, const uint4 *chi, const uint4 *clo, const float4 *ccon) {
        const int tt = T % ntiles;
        if (skips<M>() && decltype(skip_tag)::value) {
            const float tlo = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tile_tlo), tt));
            const float thi = __int_as_float(__builtin_amdgcn_readlane(__float_as_int(tile_thi), tt));
            if (tlo - wt_hi > skip_gap<M>() || wt_lo - thi > skip_gap<M>()) {  // the tile's weights vanish for these rows
                if (!diag) stage[T & 1][wave][lane] = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                nskip += (diag ? 1ull : 2ull) * (unsigned long long)(16 * RG * kBN);
                if (T % ntiles == ntiles - 1) end_sweep();
                return;
            }
        }
        const bool exact = row_small || (M == SRED_PEARSON && ((tile_exact >> tt) & 1ull) != 0ull);
        float4 *sw = stage[T & 1][wave];
        if (diag) {
            if (exact) sym_tile<M, KS, RG, true, false>(chi, clo, ccon, ahi, alo, R, cinit, accS, sw, lane, ec);
            else sym_tile<M, KS, RG, false, false>(chi, clo, ccon, ahi, alo, R, cinit, accS, sw, lane, ec);
        } else {
            if (exact) sym_tile<M, KS, RG, true, true>(chi, clo, ccon, ahi, alo, R, cinit, accS, sw, lane, ec);
            else sym_tile<M, KS, RG, false, true>(chi, clo, ccon, ahi, alo, R, cinit, accS, sw, lane, ec);
        }
        if (T % ntiles == ntiles - 1) end_sweep();
    };

    auto tiles = [&](auto skip_tag) {
    if constexpr (DB) {
        // step T: wait for tile T's copy, barrier, fold tile T-1's column sums,
        // start sweep if needed, issue tile T+1's copy, compute tile T
        auto step = [&](int T, const uint4 *chi, const uint4 *clo, const float4 *ccon, uint4 *nhi, uint4 *nlo,
                        float4 *ncon) {
            wait_vm_all();
            __syncthreads();
            if (!diag && T > 0) fold(T - 1);
            if (T % ntiles == 0) {
                // the A fragments land before the next copy is issued, so no
                // later wait for them has to count (and drain) that copy
                begin_sweep(T);
                wait_vm_all();
            }
            if (T + 1 < NT) tile_to_lds<KS>(Uhi, Ulo, aux, col0 + ((T + 1) % ntiles) * kBN, nhi, nlo, ncon, wave, lane);
            compute(skip_tag, T, chi, clo, ccon);
        };
        tile_to_lds<KS>(Uhi, Ulo, aux, col0, hi0, lo0, con0, wave, lane);
        for (int T = 0; T < NT; T += 2) {
            step(T, hi0, lo0, con0, hi1, lo1, con1);
            step(T + 1, hi1, lo1, con1, hi0, lo0, con0);
        }
    } else {
        for (int T = 0; T < NT; T++) {
            wait_vm_all();
            __syncthreads();
            if (!diag && T > 0) fold(T - 1);
            if (T % ntiles == 0) begin_sweep(T);
            tile_to_lds<KS>(Uhi, Ulo, aux, col0 + (T % ntiles) * kBN, hi0, lo0, con0, wave, lane);
            wait_vm_all();
            __syncthreads();
            compute(skip_tag, T, hi0, lo0, con0);
        }
    }
    };
    // the largest t difference of the panel pair (rows and columns are each
    // in ascending t)
    const float span = fmaxf(aux[col0 + P - 1].x - aux[row0].x, aux[row0 + P - 1].x - aux[col0].x);
    if (skips<M>() && span > skip_gap<M>()) tiles(std::true_type{});
    else tiles(std::false_type{});
    if (!diag) {
        wait_vm_all();
        __syncthreads();
        fold(NT - 1);
        wait_vm_all();
        __syncthreads();
        for (int i = tid; i < P; i += 256) pcol[(size_t)p * ldc + (col0 - coff) + i] = colacc[i];
    }
    if (nskip && lane == 0) atomicAdd(next + 1, nskip);
  }
}

// ---------------------------------------------------------------------------
// 3b. plain FP32 contraction (vector ALU).  64x64 pair tiles, 4x4 pairs per
// thread; every k step is two broadcast-free float4 LDS reads and 16 FMAs.
// The block vectors (k-major) are staged f32_chunk(K4) k values at a time:
// up to K4 = 160 (80 KiB of LDS, two workgroups per CU) one chunk holds the
// whole block and the workgroup's 64 rows stay in LDS for the whole column
// sweep; larger blocks (any K up to the footprint cap) go in chunks of 128
// and reload the row chunk with each column chunk.
// ---------------------------------------------------------------------------
__host__ __device__ constexpr int f32_chunk(int K4) { return K4 <= 160 ? K4 : 128; }

template <int M>
__global__ __launch_bounds__(256) void k_contract_f32(const float *__restrict__ UT, const float4 *__restrict__ aux,
                                                      const int *__restrict__ tile_small, int K4, int Nv_pad,
                                                      int nrowtiles, int cols_per_chunk, float *__restrict__ partial,
                                                      EpiConst ec)
{
    extern __shared__ float sm[];
    const int KC = f32_chunk(K4);
    const int nkc = (K4 + KC - 1) / KC;
    float *As = sm;
    float *Bs = sm + (size_t)KC * 64;
    const int rt = blockIdx.x % nrowtiles, ch = blockIdx.x / nrowtiles;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int row0 = rt * 64;
    // k rows [k0, k0 + kn) of 64 blocks starting at v0 into dst[k - k0][64]
    auto stage = [&](float *dst, int v0, int k0, int kn) {
        for (int i = tid; i < kn * 16; i += 256) {
            int k = i >> 4, q = i & 15;
            *reinterpret_cast<float4 *>(&dst[k * 64 + q * 4]) =
                *reinterpret_cast<const float4 *>(&UT[(size_t)(k0 + k) * Nv_pad + v0 + q * 4]);
        }
    };
    if (nkc == 1) stage(As, row0, 0, K4);
    float4 R[4];
#pragma unroll
    for (int i = 0; i < 4; i++) R[i] = row_const<M>(aux[row0 + ty * 4 + i], ec);
    const bool row_small = (M == SRED_PEARSON) && (tile_small[row0 / 32] | tile_small[row0 / 32 + 1]) != 0;
    double accS[4] = {0, 0, 0, 0};
    const int c_begin = ch * cols_per_chunk;
    const int c_end = min(c_begin + cols_per_chunk, Nv_pad);
    for (int c0 = c_begin; c0 < c_end; c0 += 64) {
        const bool exact = row_small || (M == SRED_PEARSON && (tile_small[c0 / 32] | tile_small[c0 / 32 + 1]) != 0);
        float d[4][4];
#pragma unroll
        for (int i = 0; i < 4; i++)
#pragma unroll
            for (int j = 0; j < 4; j++) d[i][j] = 0.0f;
        for (int k0 = 0; k0 < K4; k0 += KC) {
            const int kn = min(KC, K4 - k0);
            __syncthreads();
            if (nkc > 1) stage(As, row0, k0, kn);
            stage(Bs, c0, k0, kn);
            __syncthreads();
#pragma unroll 4
            for (int k = 0; k < kn; k++) {
                float4 a = *reinterpret_cast<const float4 *>(&As[k * 64 + ty * 4]);
                float4 b = *reinterpret_cast<const float4 *>(&Bs[k * 64 + tx * 4]);
                float av[4] = {a.x, a.y, a.z, a.w}, bv[4] = {b.x, b.y, b.z, b.w};
#pragma unroll
                for (int i = 0; i < 4; i++)
#pragma unroll
                    for (int j = 0; j < 4; j++) d[i][j] = fmaf(av[i], bv[j], d[i][j]);
            }
        }
        float4 C[4];
#pragma unroll
        for (int j = 0; j < 4; j++) C[j] = col_const<M>(aux[c0 + tx * 4 + j], ec);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float s = 0.0f;
            if (exact) {
#pragma unroll
                for (int j = 0; j < 4; j++) epilogue<M, true>(d[i][j], R[i], C[j], s, ec);
            } else {
#pragma unroll
                for (int j = 0; j < 4; j++) epilogue<M, false>(d[i][j], R[i], C[j], s, ec);
            }
            accS[i] += (double)s;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double s = accS[i];
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) s += __shfl_xor(s, m);
        if (tx == 0) partial[(size_t)ch * Nv_pad + row0 + ty * 4 + i] = (float)s;
    }
}

// ---------------------------------------------------------------------------
// 3c. AbsDiffStds (kernelGetGlobalAbsDiffStds2D.cl): no block contraction,
// only the weight and |s_i - s_j| per pair.  64x64 pair tiles, 4x4 per thread.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_contract_absdiff(const float4 *__restrict__ aux, int Nv_pad, int nrowtiles,
                                                          int cols_per_chunk, float *__restrict__ partial,
                                                          EpiConst ec)
{
    const int rt = blockIdx.x % nrowtiles, ch = blockIdx.x / nrowtiles;
    const int tid = threadIdx.x, tx = tid & 15, ty = tid >> 4;
    const int row0 = rt * 64;
    __shared__ float tcol[64];
    float4 R[4];
#pragma unroll
    for (int i = 0; i < 4; i++) R[i] = aux[row0 + ty * 4 + i];
    double accS[4] = {0, 0, 0, 0};
    const int c_begin = ch * cols_per_chunk;
    const int c_end = min(c_begin + cols_per_chunk, Nv_pad);
    for (int c0 = c_begin; c0 < c_end; c0 += 64) {
        __syncthreads();
        if (tid < 64) tcol[tid] = aux[c0 + tid].x;
        __syncthreads();
        float4 C[4];
#pragma unroll
        for (int j = 0; j < 4; j++) C[j] = make_float4(tcol[tx * 4 + j], 0.0f, 0.0f, 0.0f);
#pragma unroll
        for (int i = 0; i < 4; i++) {
            float s = 0.0f;
#pragma unroll
            for (int j = 0; j < 4; j++) epilogue<SRED_ABSDIFF>(0.0f, R[i], C[j], s, ec);
            accS[i] += (double)s;
        }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
        double s = accS[i];
#pragma unroll
        for (int m = 1; m < 16; m <<= 1) s += __shfl_xor(s, m);
        if (tx == 0) partial[(size_t)ch * Nv_pad + row0 + ty * 4 + i] = (float)s;
    }
}

// ---------------------------------------------------------------------------
// 4. reduce: FP64 sum of the per-chunk partials, scattered to pixel maps.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_reduce(const float *__restrict__ partial, int nchunks, int Nv, int Nv_pad,
                                                const int *__restrict__ idx, const float *__restrict__ stds,
                                                int flat_rows, float flat_count, double s_scale,
                                                float *__restrict__ S_map)
{
    int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Nv) return;
    double s = 0.0;
    for (int c = 0; c < nchunks; c++) s += (double)partial[(size_t)c * Nv_pad + v];
    int p = idx[v];
    // both-flat pairs contribute 1*weight with weight exp(-0) = 1
    // (kernelGetGlobalPearson2D.cl:101-102, kernelGetGlobalSsim2D.cl:108-109);
    // the 3D kernels skip flat rows (flat_count 0, every metric)
    if (flat_rows && stds[p] == 0.0f) s = (double)flat_count;
    S_map[p] = (float)(s * s_scale);
}

// Row totals of a contraction run in super-panels: acc[v] += the nslots
// partials of row v (slot stride ld), in slot order, in FP64.
__global__ __launch_bounds__(256) void k_accum_partials(const float *__restrict__ part, int nslots, int ld, int n,
                                                        double *__restrict__ acc)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= n) return;
    double s = 0.0;
    for (int c = 0; c < nslots; c++) s += (double)part[(size_t)c * ld + v];
    acc[v] += s;
}

// k_reduce for FP64 row totals
__global__ __launch_bounds__(256) void k_reduce_acc64(const double *__restrict__ acc, int Nv,
                                                      const int *__restrict__ idx, const float *__restrict__ stds,
                                                      int flat_rows, float flat_count, double s_scale,
                                                      float *__restrict__ S_map)
{
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Nv) return;
    double s = acc[v];
    const int p = idx[v];
    if (flat_rows && stds[p] == 0.0f) s = (double)flat_count;
    S_map[p] = (float)(s * s_scale);
}

// ---------------------------------------------------------------------------
// 5. host finalize of CLUtils.getGlobal*2D (CLUtils.java:1747-1759 and the
// inversion of :1997-2010 / :2126-2139): rep = S/(Wsum*nPixels + eps) on the
// interior in Java float order; border pixels stay 0.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_finalize(const float *__restrict__ S, const float *__restrict__ W,
                                                  const float *__restrict__ mask, int w, int h, int brw, int brh,
                                                  int d, int brd, float n_pixels, int invert, float *__restrict__ out)
{
#pragma clang fp contract(off)
    int64_t n = (int64_t)w * h * d;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t t = i / w;
        const int x = (int)(i - t * w), y = (int)(t % h), z = (int)(t / h);
        float r = 0.0f;
        if (x >= brw && x < w - brw && y >= brh && y < h - brh && z >= brd && z < d - brd) {
            float den = W[i] * n_pixels;
            den = den + kEps;
            r = S[i] / den;
            if (invert && mask[i] > 0.0f) r = (r == 0.0f) ? 1.0f : 1.0f / r;
        }
        out[i] = r;
    }
}

// ---------------------------------------------------------------------------
// host side
// ---------------------------------------------------------------------------
template <int M, int KS>
static void launch_mfma_ks(int grid, hipStream_t st, const uint4 *hi, const uint4 *lo, const float4 *aux,
                           const int *tsm, int nrt, int cpc, int ncols, float *part, int Nv_pad, EpiConst ec)
{
    hipLaunchKernelGGL((k_contract_mfma<M, KS>), dim3(grid), dim3(256), 0, st, hi, lo, aux, tsm, nrt, cpc, ncols,
                       part, Nv_pad, ec);
}

// Persistent launch over nitems panel pairs of the problems in d_probs:
// a few workgroups per CU claim items in order from a counter (zeroed here
// by a kernel, so the launch can sit in a graph); workgroups beyond the
// resident ones find the counter exhausted and exit.
template <int M, int KS>
static int launch_sym_ks(long long nitems, size_t lds, hipStream_t st, const SymProblem *d_probs, int nprob,
                         unsigned long long *d_next, int P, EpiConst ec, bool rect)
{
    int dev = 0, cus = 0;
    SRED_HIP_CHECK(hipGetDevice(&dev));
    SRED_HIP_CHECK(hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev));
    const int grid = (int)std::min<long long>(nitems, 8LL * std::max(cus, 1));
    if (grid <= 0) return SRED_OK;
    hipLaunchKernelGGL(k_zero, dim3(1), dim3(64), 0, st, reinterpret_cast<unsigned int *>(d_next), (int64_t)2);
    if (rect)
        hipLaunchKernelGGL((k_contract_mfma_sym<M, KS, true>), dim3(grid), dim3(256), lds, st, d_probs, nprob, nitems,
                           d_next, P, ec);
    else
        hipLaunchKernelGGL((k_contract_mfma_sym<M, KS, false>), dim3(grid), dim3(256), lds, st, d_probs, nprob, nitems,
                           d_next, P, ec);
    return SRED_OK;
}

template <int M>
static int launch_sym(int KS, long long nitems, size_t lds, hipStream_t st, const SymProblem *d_probs, int nprob,
                      unsigned long long *d_next, int P, EpiConst ec, bool rect)
{
    switch (KS) {
    case 1: return launch_sym_ks<M, 1>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    case 2: return launch_sym_ks<M, 2>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    case 3: return launch_sym_ks<M, 3>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    case 4: return launch_sym_ks<M, 4>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    case 5: return launch_sym_ks<M, 5>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    case 6: return launch_sym_ks<M, 6>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    case 7: return launch_sym_ks<M, 7>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    case 8: return launch_sym_ks<M, 8>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    case 9: return launch_sym_ks<M, 9>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    case 10: return launch_sym_ks<M, 10>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    case 11: return launch_sym_ks<M, 11>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    case 12: return launch_sym_ks<M, 12>(nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    default:
        set_error("the symmetric split-BF16 contraction supports block sizes up to K = 384 pixels");
        return SRED_ERR_ARG;
    }
}

int launch_sym_metric(int metric, int KS, long long nitems, hipStream_t st, const SymProblem *d_probs, int nprob,
                      unsigned long long *d_next, int P, EpiConst ec, bool rect)
{
    const size_t lds = sizeof(float) * (size_t)P;
    if (metric == SRED_PEARSON)
        return launch_sym<SRED_PEARSON>(KS, nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    if (metric == SRED_SSIM) return launch_sym<SRED_SSIM>(KS, nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    if (metric == SRED_NRMSE) return launch_sym<SRED_NRMSE>(KS, nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    if (metric == kNrmseWeighted)
        return launch_sym<kNrmseWeighted>(KS, nitems, lds, st, d_probs, nprob, d_next, P, ec, rect);
    set_error("the symmetric contraction takes Pearson, SSIM or NRMSE");
    return SRED_ERR_ARG;
}

template <int M>
static int launch_mfma(int KS, int grid, hipStream_t st, const uint4 *hi, const uint4 *lo, const float4 *aux,
                       const int *tsm, int nrt, int cpc, int ncols, float *part, int Nv_pad, EpiConst ec)
{
    switch (KS) {
    case 1: launch_mfma_ks<M, 1>(grid, st, hi, lo, aux, tsm, nrt, cpc, ncols, part, Nv_pad, ec); break;
    case 2: launch_mfma_ks<M, 2>(grid, st, hi, lo, aux, tsm, nrt, cpc, ncols, part, Nv_pad, ec); break;
    case 3: launch_mfma_ks<M, 3>(grid, st, hi, lo, aux, tsm, nrt, cpc, ncols, part, Nv_pad, ec); break;
    case 4: launch_mfma_ks<M, 4>(grid, st, hi, lo, aux, tsm, nrt, cpc, ncols, part, Nv_pad, ec); break;
    case 5: launch_mfma_ks<M, 5>(grid, st, hi, lo, aux, tsm, nrt, cpc, ncols, part, Nv_pad, ec); break;
    case 6: launch_mfma_ks<M, 6>(grid, st, hi, lo, aux, tsm, nrt, cpc, ncols, part, Nv_pad, ec); break;
    case 7: launch_mfma_ks<M, 7>(grid, st, hi, lo, aux, tsm, nrt, cpc, ncols, part, Nv_pad, ec); break;
    case 8: launch_mfma_ks<M, 8>(grid, st, hi, lo, aux, tsm, nrt, cpc, ncols, part, Nv_pad, ec); break;
    default:
        set_error("split-BF16 contraction supports block sizes up to K = 256 pixels");
        return SRED_ERR_ARG;
    }
    return SRED_OK;
}

template <int M>
static void launch_f32(int grid, size_t lds, hipStream_t st, const float *UT, const float4 *aux, const int *tsm,
                       int K4, int Nv_pad, int nrt, int cpc, float *part, EpiConst ec)
{
    hipLaunchKernelGGL((k_contract_f32<M>), dim3(grid), dim3(256), lds, st, UT, aux, tsm, K4, Nv_pad, nrt, cpc, part,
                       ec);
}

// Columns per chunk: the chunk's column data (~2 MiB) stays resident in an
// XCD's L2 while the concurrently running workgroups sweep it; smaller chunks
// when there are too few row tiles to fill 256 CUs.
static int choose_chunk_cols(int Nv_pad, int bytes_per_col, int nrowtiles, int override_cols, int max_cols)
{
    int cols;
    if (override_cols > 0) {
        cols = std::min((override_cols + 63) / 64 * 64, max_cols);
    } else {
        cols = std::max(64, (2 << 20) / std::max(bytes_per_col, 1) / 64 * 64);
        const int target_wgs = 4096;
        int want_chunks = (target_wgs + nrowtiles - 1) / nrowtiles;
        int cols_fill = ((Nv_pad + want_chunks - 1) / want_chunks + 63) / 64 * 64;
        cols = std::max(64, std::min(std::min(cols, cols_fill), max_cols));
    }
    return std::min(cols, Nv_pad);
}

void launch_gather_bf16(sred_ctx *ctx, int metric, const float *d_img, const float *d_means, const float *d_stds,
                        const int *d_idx, int Nv, int Nv_pad, int K, int KS, const int *d_offs, float inv_sqrt_km1,
                        float inv_sqrt_k, uint4 *Uhi, uint4 *Ulo)
{
    size_t u4 = (size_t)Nv_pad * KS * 4;
    hipLaunchKernelGGL(k_gather_bf16, dim3((unsigned)((u4 + 255) / 256)), dim3(256), 0, ctx->stream, metric, d_img,
                       d_means, d_stds, d_idx, Nv, Nv_pad, K, KS, d_offs, inv_sqrt_km1, inv_sqrt_k, 0, Uhi, Ulo);
}

void launch_gather_f32(sred_ctx *ctx, int metric, const float *d_img, const float *d_means, const float *d_stds,
                       const int *d_idx, int Nv, int Nv_pad, int K, int K4, const int *d_offs, float inv_sqrt_km1,
                       float inv_sqrt_k, float *UT)
{
    size_t nu = (size_t)Nv_pad * K4;
    hipLaunchKernelGGL(k_gather_f32, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0, ctx->stream, metric, d_img,
                       d_means, d_stds, d_idx, Nv, Nv_pad, K, K4, d_offs, inv_sqrt_km1, inv_sqrt_k, UT);
}

int zero_dev(sred_ctx *ctx, void *ptr, size_t bytes)
{
    int64_t words = (int64_t)((bytes + 3) / 4);
    int blocks = (int)std::min<int64_t>((words + 255) / 256, 4096);
    if (blocks < 1) blocks = 1;
    hipLaunchKernelGGL(k_zero, dim3(blocks), dim3(256), 0, ctx->stream, (unsigned int *)ptr, words);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

// Launch geometry of the contraction for Nv valid blocks of K pixels.
int plan_contraction(int Nv, int K, int metric, int precision, int symmetric, int chunk_cols, ContractionPlan *pl)
{
    return plan_contraction_p(Nv, K, metric, precision, symmetric, chunk_cols, 0, pl);
}

int plan_contraction_p(int Nv, int K, int metric, int precision, int symmetric, int chunk_cols, int panel,
                       ContractionPlan *pl)
{
    ContractionPlan q{};
    if (Nv < 0 || K < 2 || metric < 0 || metric > 3) {
        set_error("invalid contraction plan arguments");
        return SRED_ERR_ARG;
    }
    // split BF16 stores K - 1 reflected components when that saves a k-step
    // (K = 32m + 1; k_gather_bf16)
    q.reflect = (precision == SRED_BF16X3) && metric != SRED_ABSDIFF && K > 1 && (K - 1) % 32 == 0;
    q.KS = q.reflect ? (K - 1) / 32 : (K + 31) / 32;
    q.K4 = (K + 3) / 4 * 4;
    // the split-BF16 kernels hold a block's fragments in registers up to
    // K_pad = 32 * kMaxKS (the symmetric one, with one row group beyond 256,
    // up to 32 * kMaxKSSym); larger blocks take the FP32 contraction, which
    // stages any K through LDS in chunks
    q.mfma = (precision == SRED_BF16X3) && metric != SRED_ABSDIFF && q.KS <= (symmetric ? kMaxKSSym : kMaxKS);
    q.reflect = q.reflect && q.mfma;
    q.sym = q.mfma && symmetric;
    // symmetric panels: P rows = P columns, a multiple of 256, up to 2048
    // (1024 at K_pad = 128, whose two tile buffers leave room for a 4 KiB
    // column accumulator only), small enough for at least ~2000 panel pairs
    q.P = q.sym ? std::min(q.KS == 4 ? 1024 : 2048, std::max(256, ((Nv + 63) / 64 + 255) / 256 * 256)) : 0;
    if (q.sym && panel > 0) {
        if (panel % 256 != 0 || panel > 2048) {
            set_error("panel size must be a multiple of 256, at most 2048");
            return SRED_ERR_ARG;
        }
        q.P = panel;
    }
    q.Nv_pad = q.sym ? (Nv + q.P - 1) / q.P * q.P : (Nv + 255) / 256 * 256;
    int bytes_per_col;
    if (q.sym) { q.bm = 64 * rg_sym(metric, q.KS); bytes_per_col = q.KS * 32 * 4; }
    else if (q.mfma) { q.bm = 64 * rg_for(metric, q.KS); bytes_per_col = q.KS * 32 * 4; }
    else if (metric == SRED_ABSDIFF) { q.bm = 64; bytes_per_col = 16; }
    else { q.bm = 64; bytes_per_col = q.K4 * 4; }
    q.nrowtiles = q.Nv_pad / q.bm;
    // the MFMA kernels keep FP32 row sums over a chunk: at most 2048 columns
    // (128 terms per lane) before the FP64 sum over chunks
    q.cols_per_chunk = q.sym ? q.P
                             : choose_chunk_cols(std::max(q.Nv_pad, 64), bytes_per_col, std::max(q.nrowtiles, 1),
                                                 chunk_cols, q.mfma ? 2048 : 1 << 30);
    q.nchunks = q.Nv_pad > 0 ? (q.Nv_pad + q.cols_per_chunk - 1) / q.cols_per_chunk : 0;
    q.grid = q.sym ? (int64_t)q.nchunks * (q.nchunks + 1) / 2 : (int64_t)q.nrowtiles * q.nchunks;
    q.partial_elems = (int64_t)q.nchunks * q.Nv_pad;
    q.pairs_evaluated = q.sym ? (int64_t)q.P * q.P * q.grid : (int64_t)q.Nv_pad * q.Nv_pad;
    const bool ok = q.sym ? (q.Nv_pad % q.P == 0 && q.P % q.bm == 0 && q.nchunks * q.P == q.Nv_pad)
                          : (q.Nv_pad % q.bm == 0 && q.nrowtiles * q.bm == q.Nv_pad &&
                             (int64_t)q.nchunks * q.cols_per_chunk >= q.Nv_pad && q.cols_per_chunk % 64 == 0);
    if (!ok || q.grid > 0x7fffffff) {
        set_error("contraction geometry out of range");
        return SRED_ERR_ARG;
    }
    *pl = q;
    return SRED_OK;
}

// Buffers of one global problem: the context's own for single calls, one set
// per image for a batch.
GlobalBufs ctx_global_bufs(sred_ctx *ctx)
{
    return GlobalBufs{&ctx->counts, &ctx->valid_idx, &ctx->aux, &ctx->tflags, &ctx->u32, &ctx->uhi, &ctx->ulo,
                      &ctx->offs2, ctx->offs2_key, &ctx->wmom, &ctx->wpart, &ctx->sortbuf};
}

// Steps 1-2 of a global problem: compaction (one synchronisation, for Nv),
// weight sums into d_W (zero elsewhere, as d_S), the launch plan (panel size
// `panel` for the symmetric contraction, 0 = automatic), per-block scalars,
// small-sigma flags and the gathered block matrix.  pad_rows > 0 pads the
// split-BF16 block matrix and scalars to a multiple of pad_rows rows, so any
// panel size dividing it can be used afterwards.
// Step 1 of a global problem, enqueued only: the compaction of the valid
// blocks, and a copy of its totals (valid and flat counts, the bits of the
// largest valid sigma) into h_tot[2] (host memory; read after the stream has
// been synchronised).
int global_prepare_front(sred_ctx *ctx, const GlobalBufs &b, const Geometry &g, const GlobalParams &p,
                         const float *d_stds, int2 *h_tot)
{
    if (p.metric < 0 || p.metric > 3) {
        set_error("unknown metric");
        return SRED_ERR_UNKNOWN_METRIC;
    }
    const double thr_d = java_float_literal(p.threshold);
    const int64_t n = g.n();
    const int nblk = (int)((n + kCompactTile - 1) / kCompactTile);
    SRED_TRY(b.counts->ensure(sizeof(int2) * (3 * (size_t)nblk + 4)));
    int2 *d_counts = b.counts->as<int2>();
    int2 *d_offsets = d_counts + nblk;                                   // nblk + 2 entries
    float *d_smax = reinterpret_cast<float *>(d_offsets + nblk + 2);  // nblk entries
    SRED_TRY(b.valid_idx->ensure(sizeof(int) * (size_t)n + 256));
    int *d_idx = b.valid_idx->as<int>();
    hipLaunchKernelGGL(k_compact_count, dim3(nblk), dim3(kCompactBlock), 0, ctx->stream, d_stds, g.w, g.h, g.brw,
                       g.brh, g.d, g.brd, thr_d, p.drop_flat ? 1 : 0, d_counts, d_smax);
    hipLaunchKernelGGL(k_compact_scan, dim3(1), dim3(1024), 0, ctx->stream, d_counts, d_smax, nblk, d_offsets);
    hipLaunchKernelGGL(k_compact_scatter, dim3(nblk), dim3(kCompactBlock), 0, ctx->stream, d_stds, g.w, g.h, g.brw,
                       g.brh, g.d, g.brd, thr_d, p.drop_flat ? 1 : 0, d_offsets, d_idx);
    SRED_HIP_CHECK(hipGetLastError());
    SRED_HIP_CHECK(hipMemcpyAsync(h_tot, d_offsets + nblk, 2 * sizeof(int2), hipMemcpyDeviceToHost, ctx->stream));
    return SRED_OK;
}

int global_prepare(sred_ctx *ctx, const GlobalBufs &b, const Geometry &g, const GlobalParams &p, const float *d_img,
                   const float *d_means, const float *d_stds, float *d_S, float *d_W, int symmetric, int panel,
                   int pad_rows, GlobalPrepared *out)
{
    int2 tot[2];
    SRED_TRY(global_prepare_front(ctx, b, g, p, d_stds, tot));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return global_prepare_back(ctx, b, g, p, d_img, d_means, d_stds, d_S, d_W, symmetric, panel, pad_rows, tot, out);
}

// Steps 1-2 after the compaction totals tot[2] have reached the host.
int global_prepare_back(sred_ctx *ctx, const GlobalBufs &b, const Geometry &g, const GlobalParams &p,
                        const float *d_img, const float *d_means, const float *d_stds, float *d_S, float *d_W,
                        int symmetric, int panel, int pad_rows, const int2 *tot, GlobalPrepared *out)
{
    const int metric = p.metric;
    GlobalPrepared r{};
    const double fp_d = java_float_literal(p.filter_param);
    const double eps_d = java_float_literal(kEps);
    r.tscale_d = std::sqrt(1.4426950408889634 / (fp_d + eps_d));
    const float tscale = (float)r.tscale_d;
    const float inv_sqrt_km1 = (float)(1.0 / std::sqrt((double)(g.K - 1)));
    const float inv_sqrt_k = (float)(1.0 / std::sqrt((double)g.K));
    const int64_t n = g.n();
    int *d_idx = b.valid_idx->as<int>();
    const int Nv = tot[0].x;
    float sigma_max;
    std::memcpy(&sigma_max, &tot[1].x, sizeof sigma_max);
    ctx->last_sigma_max = sigma_max;
    r.Nv = Nv;
    // both-flat pairs give a flat row the flat-column count; rows the 3D kernels skip score 0
    r.flat_count = p.skip_flat_rows ? 0.0f : (float)std::min(tot[0].y, 16777216);
    r.zero_flat_rows = p.skip_flat_rows;
    r.d_idx = d_idx;

    SRED_TRY(zero_dev(ctx, d_S, sizeof(float) * n));
    SRED_TRY(zero_dev(ctx, d_W, sizeof(float) * n));
    if (Nv == 0) {
        *out = r;
        return SRED_OK;
    }
    // weight sums (independent of the contraction)
    SRED_TRY(weight_sums_dev(ctx, *b.wmom, *b.wpart, d_stds, d_idx, Nv, r.tscale_d, sigma_max, d_W));
    SRED_TRY(plan_contraction_p(Nv, g.K, metric, p.precision, symmetric, ctx->chunk_cols, panel, &r.pl));
    const ContractionPlan &pl = r.pl;
    const bool mfma = pl.mfma;
    const int Nv_rows = (pad_rows > 0 && mfma) ? std::max(pl.Nv_pad, (Nv + pad_rows - 1) / pad_rows * pad_rows)
                                               : pl.Nv_pad;
    r.Nv_rows = Nv_rows;
    if (mfma && Nv > 1) {
        // contraction order: ascending sigma (lets the symmetric kernel skip
        // panel pairs and tiles whose weights vanish)
        int *sorted = nullptr;
        SRED_TRY(sort_valid_by_sigma(ctx, *b.sortbuf, d_stds, d_idx, Nv, &sorted));
        d_idx = sorted;
        r.d_idx = d_idx;
    }

    // 2. offsets relative to the image (or volume) and per-block scalars
    SRED_TRY(upload_offsets(ctx, g, g.w, b.offs2, b.offs2_key));
    const int *d_offs = b.offs2->as<int>();
    SRED_TRY(b.aux->ensure(sizeof(float4) * (size_t)Nv_rows));
    float4 *d_aux = b.aux->as<float4>();
    hipLaunchKernelGGL(k_aux, dim3((Nv_rows + 255) / 256), dim3(256), 0, ctx->stream, metric, d_img, d_means, d_stds,
                       d_idx, Nv, Nv_rows, g.K, d_offs, tscale, inv_sqrt_km1, inv_sqrt_k, mfma ? 1 : 0,
                       pl.reflect ? 1 : 0, d_aux);
    const int ngroups = Nv_rows / 32;
    SRED_TRY(b.tflags->ensure(sizeof(int) * (size_t)std::max(ngroups, 1)));
    hipLaunchKernelGGL(k_tile_small, dim3((ngroups * 32 + 255) / 256), dim3(256), 0, ctx->stream, d_aux, ngroups,
                       b.tflags->as<int>());
    r.d_aux = d_aux;
    r.d_tsm = b.tflags->as<int>();
    if (mfma) {
        size_t u4 = (size_t)Nv_rows * pl.KS * 4;
        SRED_TRY(b.uhi->ensure(sizeof(uint4) * u4));
        SRED_TRY(b.ulo->ensure(sizeof(uint4) * u4));
        hipLaunchKernelGGL(k_gather_bf16, dim3((unsigned)((u4 + 255) / 256)), dim3(256), 0, ctx->stream, metric, d_img,
                           d_means, d_stds, d_idx, Nv, Nv_rows, g.K, pl.KS, d_offs, inv_sqrt_km1, inv_sqrt_k,
                           pl.reflect ? 1 : 0, b.uhi->as<uint4>(), b.ulo->as<uint4>());
        r.uhi = b.uhi->as<uint4>();
        r.ulo = b.ulo->as<uint4>();
    } else if (metric != SRED_ABSDIFF) {
        size_t nu = (size_t)pl.Nv_pad * pl.K4;
        SRED_TRY(b.u32->ensure(sizeof(float) * nu));
        hipLaunchKernelGGL(k_gather_f32, dim3((unsigned)((nu + 255) / 256)), dim3(256), 0, ctx->stream, metric, d_img,
                           d_means, d_stds, d_idx, Nv, pl.Nv_pad, g.K, pl.K4, d_offs, inv_sqrt_km1, inv_sqrt_k,
                           b.u32->as<float>());
        r.u32 = b.u32->as<float>();
    }
    SRED_HIP_CHECK(hipGetLastError());
    *out = r;
    return SRED_OK;
}

// Step 4: FP64 sum of a problem's partials (nchunks slots of stride Nv_pad)
// into the S map; flat rows of Pearson/SSIM get the flat-column count.
int global_reduce_dev(sred_ctx *ctx, const GlobalPrepared &r, int metric, const float *d_part, int nchunks,
                      int Nv_pad, const float *d_stds, float *d_S)
{
    if (r.Nv == 0) return SRED_OK;
    const double s_scale = (metric == SRED_ABSDIFF) ? 1.0 / (double)(float)r.tscale_d : 1.0;
    const int flat_rows = r.zero_flat_rows || metric == SRED_PEARSON || metric == SRED_SSIM;
    hipLaunchKernelGGL(k_reduce, dim3((r.Nv + 255) / 256), dim3(256), 0, ctx->stream, d_part, nchunks, r.Nv, Nv_pad,
                       r.d_idx, d_stds, flat_rows, r.flat_count, s_scale, d_S);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

EpiConst epi_const() { return EpiConst{kEps, 0.0001f, 0.0009f, (float)((double)kEps * 1.4426950408889634)}; }

// The symmetric contraction of a call whose partial sums (npanels x Nv_pad
// floats) exceed `budget` floats: the blocks (in sigma order) are cut into
// super-panels of S blocks, S the largest multiple of the panel size P with
// 2 (S/P) S floats within the budget (at least P), and each super-panel's
// triangle and each pair of super-panels' rectangle is one persistent launch
// over the same panel pairs the whole call would evaluate.  After each launch
// its partials are added, slot by slot in FP64, to the row totals of its
// blocks (rows and, for a rectangle, columns); the launches run in a fixed
// order, so the result is deterministic.  The partial-sum buffer is reused.
static int sym_superpanels(sred_ctx *ctx, const GlobalPrepared &r, int cm, int metric, const EpiConst &ec,
                           int64_t budget, const float *d_stds, float *d_S)
{
    const ContractionPlan &pl = r.pl;
    const int P = pl.P, KS = pl.KS, Nv_pad = pl.Nv_pad;
    int64_t nps = 1;  // panels per super-panel: 2 (S/P) S = 2 nps^2 P floats of partials
    while (2 * (nps + 1) * (nps + 1) * (int64_t)P <= budget) nps++;
    const int S = (int)std::min<int64_t>(nps * P, Nv_pad);
    const int m = (Nv_pad + S - 1) / S;
    SRED_TRY(ctx->partial.ensure(sizeof(float) * 2 * (size_t)(S / P) * S));
    SRED_TRY(ctx->acc64.ensure(sizeof(double) * (size_t)Nv_pad));
    float *d_part = ctx->partial.as<float>();
    double *d_acc = ctx->acc64.as<double>();
    SRED_TRY(zero_dev(ctx, d_acc, sizeof(double) * (size_t)Nv_pad));
    // every super-panel pair's table entry, uploaded once (the table outlives the launches)
    ctx->host_symtab.clear();
    std::vector<long long> nitems;
    for (int a = 0; a < m; a++)
        for (int b = a; b < m; b++) {
            const int A0 = a * S, SA = std::min(S, Nv_pad - A0), B0 = b * S, SB = std::min(S, Nv_pad - B0);
            const size_t fo = (size_t)(A0 / 16) * KS * 64;
            SymProblem q{r.uhi + fo, r.ulo + fo, r.d_aux + A0, r.d_tsm + A0 / 32, d_part, SA, SA / P, 0};
            if (a != b) {
                q.pcol = d_part + (size_t)(SB / P) * SA;
                q.ldc = SB;
                q.coff = B0 - A0;
                q.npr = SA / P;
                nitems.push_back((long long)(SA / P) * (SB / P));
            } else {
                nitems.push_back((long long)(SA / P) * (SA / P + 1) / 2);
            }
            ctx->host_symtab.push_back(q);
        }
    const size_t ntab = ctx->host_symtab.size();
    const size_t tab_bytes = (sizeof(SymProblem) * ntab + 127) / 128 * 128;
    SRED_TRY(ctx->symtab.ensure(tab_bytes + 256));
    SRED_HIP_CHECK(hipMemcpyAsync(ctx->symtab.ptr, ctx->host_symtab.data(), sizeof(SymProblem) * ntab,
                                  hipMemcpyHostToDevice, ctx->stream));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const SymProblem *d_tab = ctx->symtab.as<SymProblem>();
    unsigned long long *d_next = reinterpret_cast<unsigned long long *>(ctx->symtab.as<char>() + tab_bytes);
    SRED_TRY(zero_dev(ctx, d_next + 1, sizeof(unsigned long long)));
    ctx->last_skip = d_next + 1;
    SRED_TRY(time_begin(ctx));
    size_t t = 0;
    for (int a = 0; a < m; a++)
        for (int b = a; b < m; b++, t++) {
            const int A0 = a * S, SA = std::min(S, Nv_pad - A0), B0 = b * S, SB = std::min(S, Nv_pad - B0);
            SRED_TRY(launch_sym_metric(cm, KS, nitems[t], ctx->stream, d_tab + t, 1, d_next, P, ec, true));
            SRED_HIP_CHECK(hipGetLastError());
            const int nrow_slots = (a == b) ? SA / P : SB / P;
            hipLaunchKernelGGL(k_accum_partials, dim3((SA + 255) / 256), dim3(256), 0, ctx->stream, d_part,
                               nrow_slots, SA, SA, d_acc + A0);
            if (a != b)
                hipLaunchKernelGGL(k_accum_partials, dim3((SB + 255) / 256), dim3(256), 0, ctx->stream,
                                   d_part + (size_t)(SB / P) * SA, SA / P, SB, SB, d_acc + B0);
            SRED_HIP_CHECK(hipGetLastError());
        }
    SRED_TRY(time_end(ctx, pl.pairs_evaluated));
    ctx->last_launches = (int)ntab;
    const double s_scale = (metric == SRED_ABSDIFF) ? 1.0 / (double)(float)r.tscale_d : 1.0;
    const int flat_rows = r.zero_flat_rows || metric == SRED_PEARSON || metric == SRED_SSIM;
    hipLaunchKernelGGL(k_reduce_acc64, dim3((r.Nv + 255) / 256), dim3(256), 0, ctx->stream, d_acc, r.Nv, r.d_idx,
                       d_stds, flat_rows, r.flat_count, s_scale, d_S);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

int global_accum_dev(sred_ctx *ctx, const Geometry &g, const GlobalParams &p, const float *d_img,
                     const float *d_means, const float *d_stds, float *d_S, float *d_W)
{
    const int metric = p.metric;
    // template id of the contraction (data layout and reduction follow `metric`)
    const int cm = (metric == SRED_NRMSE && p.nrmse_weighted) ? kNrmseWeighted : metric;
    GlobalPrepared r;
    ctx->last_ms = 0.0f;
    ctx->last_launches = 0;
    ctx->last_skip = nullptr;
    SRED_TRY(global_prepare(ctx, ctx_global_bufs(ctx), g, p, d_img, d_means, d_stds, d_S, d_W, ctx->symmetric, 0, 0,
                            &r));
    ctx->last_nv = r.Nv;
    ctx->last_pairs = (int64_t)r.Nv * r.Nv;
    if (r.Nv == 0) return SRED_OK;
    const ContractionPlan &pl = r.pl;
    const int KS = pl.KS, K4 = pl.K4, P = pl.P, Nv_pad = pl.Nv_pad;
    const int nrowtiles = pl.nrowtiles, cpc = pl.cols_per_chunk, nchunks = pl.nchunks;
    const int grid = (int)pl.grid;
    const EpiConst ec = epi_const();
    const int64_t budget = ctx->partial_mb << 18;  // floats
    if (pl.sym && pl.partial_elems > budget) return sym_superpanels(ctx, r, cm, metric, ec, budget, d_stds, d_S);
    SRED_TRY(ctx->partial.ensure(sizeof(float) * (size_t)pl.partial_elems));
    float *d_part = ctx->partial.as<float>();

    // 3. contraction (timed with events on the launch stream)
    SRED_TRY(time_begin(ctx));
    if (pl.sym) {
        // one problem: its table entry lives in the context (the copy reads it
        // before this call returns: the stream was synchronised above)
        ctx->host_symtab.assign(1, SymProblem{r.uhi, r.ulo, r.d_aux, r.d_tsm, d_part, Nv_pad, Nv_pad / P, 0});
        SRED_TRY(ctx->symtab.ensure(sizeof(SymProblem) + 256));
        SRED_HIP_CHECK(hipMemcpyAsync(ctx->symtab.ptr, ctx->host_symtab.data(), sizeof(SymProblem),
                                      hipMemcpyHostToDevice, ctx->stream));
        unsigned long long *d_next =
            reinterpret_cast<unsigned long long *>(ctx->symtab.as<char>() + (sizeof(SymProblem) + 127) / 128 * 128);
        SRED_TRY(zero_dev(ctx, d_next + 1, sizeof(unsigned long long)));
        ctx->last_skip = d_next + 1;
        SRED_TRY(launch_sym_metric(cm, KS, pl.grid, ctx->stream, ctx->symtab.as<SymProblem>(), 1, d_next, P, ec));
    } else if (pl.mfma) {
        int rc = SRED_OK;
        switch (cm) {
        case SRED_PEARSON:
            rc = launch_mfma<SRED_PEARSON>(KS, grid, ctx->stream, r.uhi, r.ulo, r.d_aux, r.d_tsm, nrowtiles, cpc,
                                           Nv_pad, d_part, Nv_pad, ec);
            break;
        case SRED_SSIM:
            rc = launch_mfma<SRED_SSIM>(KS, grid, ctx->stream, r.uhi, r.ulo, r.d_aux, r.d_tsm, nrowtiles, cpc, Nv_pad,
                                        d_part, Nv_pad, ec);
            break;
        case kNrmseWeighted:
            rc = launch_mfma<kNrmseWeighted>(KS, grid, ctx->stream, r.uhi, r.ulo, r.d_aux, r.d_tsm, nrowtiles, cpc,
                                             Nv_pad, d_part, Nv_pad, ec);
            break;
        default:
            rc = launch_mfma<SRED_NRMSE>(KS, grid, ctx->stream, r.uhi, r.ulo, r.d_aux, r.d_tsm, nrowtiles, cpc,
                                         Nv_pad, d_part, Nv_pad, ec);
            break;
        }
        SRED_TRY(rc);
    } else if (metric == SRED_ABSDIFF) {
        hipLaunchKernelGGL(k_contract_absdiff, dim3(grid), dim3(256), 0, ctx->stream, r.d_aux, Nv_pad, nrowtiles, cpc,
                           d_part, ec);
    } else {
        const size_t lds = sizeof(float) * 2 * (size_t)f32_chunk(K4) * 64;
        if (metric == SRED_PEARSON)
            launch_f32<SRED_PEARSON>(grid, lds, ctx->stream, r.u32, r.d_aux, r.d_tsm, K4, Nv_pad, nrowtiles, cpc,
                                     d_part, ec);
        else if (metric == SRED_SSIM)
            launch_f32<SRED_SSIM>(grid, lds, ctx->stream, r.u32, r.d_aux, r.d_tsm, K4, Nv_pad, nrowtiles, cpc, d_part,
                                  ec);
        else if (cm == kNrmseWeighted)
            launch_f32<kNrmseWeighted>(grid, lds, ctx->stream, r.u32, r.d_aux, r.d_tsm, K4, Nv_pad, nrowtiles, cpc,
                                       d_part, ec);
        else
            launch_f32<SRED_NRMSE>(grid, lds, ctx->stream, r.u32, r.d_aux, r.d_tsm, K4, Nv_pad, nrowtiles, cpc,
                                   d_part, ec);
    }
    SRED_HIP_CHECK(hipGetLastError());
    SRED_TRY(time_end(ctx, pl.pairs_evaluated));

    // 4. reduce
    return global_reduce_dev(ctx, r, metric, d_part, nchunks, Nv_pad, d_stds, d_S);
}

int global_finalize_dev(sred_ctx *ctx, const Geometry &g, int metric, const float *d_S, const float *d_W,
                        const float *d_mask, float n_pixels, float c, int normalize, float *d_out)
{
    int blocks = (int)std::min<int64_t>((g.n() + 255) / 256, 8192);
    int invert = (metric == SRED_NRMSE || metric == SRED_ABSDIFF);
    hipLaunchKernelGGL(k_finalize, dim3(blocks), dim3(256), 0, ctx->stream, d_S, d_W, d_mask, g.w, g.h, g.brw, g.brh,
                       g.d, g.brd, n_pixels, invert, d_out);
    SRED_HIP_CHECK(hipGetLastError());
    if (c > 0.0f) SRED_TRY(apply_mask_dev(ctx, g, d_out, d_mask));
    if (normalize) SRED_TRY(normalize_map_dev(ctx, g, d_out, d_mask, 0));
    return SRED_OK;
}

}  // namespace sred
