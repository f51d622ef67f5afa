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
#include <type_traits>

#include "sred_internal.h"

namespace sred {

// Per valid block: (s, a, b, t) with t = s*sqrt(log2 e / H) for the weight.
__global__ __launch_bounds__(256) void k_adjoint_coef(const int *__restrict__ idx, int Nv, int Nv_pad,
                                                      const float *__restrict__ stds, const float *__restrict__ S,
                                                      const float *__restrict__ W, const float *__restrict__ mask,
                                                      const float *__restrict__ dmap, const unsigned int *__restrict__ mm,
                                                      float n_pixels, float tscale, float4 *__restrict__ info,
                                                      float4 *__restrict__ info2, unsigned int *__restrict__ bmax)
{
    // Nv_pad is a multiple of 64: the threads past it are whole waves
    const int v = blockIdx.x * blockDim.x + threadIdx.x;
    if (v >= Nv_pad) return;
    unsigned int mb = 0u, ma = 0u;
    if (v >= Nv) {
        info[v] = make_float4(0.0f, 0.0f, 0.0f, 1e30f);  // padding: weight exp2(-inf) = 0, a = b = 0
        info2[v] = make_float4(1e30f, 0.0f, 0.0f, 0.0f);
    } else {
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
        info2[v] = make_float4(s * tscale, (float)a, (float)b, s > 0.0f ? 1.0f / s : 0.0f);
        mb = __float_as_uint(fabsf((float)b));
        ma = __float_as_uint(fabsf((float)a));
    }
    // bmax[0] = max |b|, bmax[1] = max |a| over the blocks (bits of a non-negative
    // float order as unsigned; NaN above every finite value), one atomic per wave
#pragma unroll
    for (int o = 1; o < 64; o <<= 1) {
        mb = max(mb, (unsigned int)__shfl_xor((int)mb, o));
        ma = max(ma, (unsigned int)__shfl_xor((int)ma, o));
    }
    if ((threadIdx.x & 63) == 0) {
        atomicMax(bmax, mb);
        atomicMax(bmax + 1, ma);
    }
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

// ---------------------------------------------------------------------------
// The adjoint pass, split-BF16 on MFMA (every block the adjoint takes, up to K = 160).
//
// Per 64-column tile a wave (32 rows i) computes the transposed dot-product
// tile D[j][i] = u_j . u_i with v_mfma_f32_16x16x32_bf16 (A = the column
// blocks' fragments from LDS, B = the wave's row fragments in registers;
// hi*hi + hi*lo + lo*hi), then per pair the coefficient P_ij and the
// dL/ds_i term on the vector ALU.  In the MFMA output layout lane l holds
// i = l % 16 and j = 4(l / 16) + r, which is already the B-operand layout of
// the second product G^T[k][i] = sum_j u_jk P_ij when the column blocks'
// transposed fragments use the same j assignment: V[jg][kg][lane] holds
// u_{j,k}, k = 16kg + l % 16, j = 32jg + jmap(l / 16, e), jmap(q, e) = 4q + e
// for e < 4 and 16 + 4q + e - 4 beyond.  P is split into hi + lo BF16 and the
// second product uses three MFMAs as well; its FP32 accumulators are flushed
// into FP32 totals every 128 tiles, dL/ds_i into FP64 every tile.
// ---------------------------------------------------------------------------
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float floatx2 __attribute__((ext_vector_type(2)));

// hi = RNE bf16 pair of (a, b) (one v_cvt_pk_bf16_f32), lo = the same of the remainders
// The BF16 pairs (-1, 0) and (0, -1) as 32-bit words in SGPRs, for
// split_bf16x2_dot below.  They are set by s_mov_b32 in inline asm so that the
// compiler cannot fold them: it encodes the pair (-1, 0) (word 0x0000bf80)
// as the 32-bit inline constant -1.0 (word 0xbf800000), which
// v_dot2c_f32_bf16 reads as the pair (0, -1).
struct Bf16Neg {
    bf16x2 x, y;
};
__device__ __forceinline__ Bf16Neg bf16_neg_units()
{
    unsigned int kx, ky;
    asm volatile("s_mov_b32 %0, 0xbf80" : "=s"(kx));
    asm volatile("s_mov_b32 %0, 0xbf800000" : "=s"(ky));
    return Bf16Neg{__builtin_bit_cast(bf16x2, kx), __builtin_bit_cast(bf16x2, ky)};
}
// (a, b) -> BF16 pairs hi = rne(a, b), lo = rne(a - hi.x, b - hi.y), the
// residuals taken by v_dot2_f32_bf16 (hi . (-1, 0) + a): one instruction
// each instead of a widening shift or mask and a subtraction.  The residual
// of a round-to-nearest BF16 is exactly representable in FP32, so this gives
// the bits of the subtraction form.
__device__ __forceinline__ void split_bf16x2_dot(float a, float b, const Bf16Neg &n, unsigned int &hi,
                                                 unsigned int &lo)
{
    const bf16x2 h = __builtin_convertvector(floatx2{a, b}, bf16x2);
    const float la = __builtin_amdgcn_fdot2_f32_bf16(h, n.x, a, false);
    const float lb = __builtin_amdgcn_fdot2_f32_bf16(h, n.y, b, false);
    const bf16x2 l = __builtin_convertvector(floatx2{la, lb}, bf16x2);
    hi = __builtin_bit_cast(unsigned int, h);
    lo = __builtin_bit_cast(unsigned int, l);
}

// V for the k-groups kg < KG (16 k values each; KG <= 2 KS).
__global__ __launch_bounds__(256) void k_transpose_frag(const uint4 *__restrict__ Uhi, const uint4 *__restrict__ Ulo,
                                                        int Nv_pad, int KS, int KG, uint4 *__restrict__ Vhi,
                                                        uint4 *__restrict__ Vlo)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (int64_t)(Nv_pad / 32) * KG * 64) return;
    const int lane = (int)(t & 63);
    const int64_t rr = t >> 6;
    const int kg = (int)(rr % KG), jg = (int)(rr / KG);
    const int k = 16 * kg + (lane & 15), q = lane >> 4;
    // u_{j,k} sits in U[j/16][k/32][(j%16) + 16((k%32)/8)], element k%8
    const int ks = k >> 5, lq = (k & 31) >> 3, e8 = k & 7;
    unsigned int h[4] = {0u, 0u, 0u, 0u}, l[4] = {0u, 0u, 0u, 0u};
#pragma unroll
    for (int e = 0; e < 8; e++) {
        const int j = 32 * jg + (e < 4 ? 4 * q + e : 16 + 4 * q + (e - 4));
        const size_t o = ((size_t)(j >> 4) * KS + ks) * 64 + (j & 15) + 16 * lq;
        const unsigned int hw = reinterpret_cast<const unsigned short *>(&Uhi[o])[e8];
        const unsigned int lw = reinterpret_cast<const unsigned short *>(&Ulo[o])[e8];
        h[e >> 1] |= (e & 1) ? hw << 16 : hw;
        l[e >> 1] |= (e & 1) ? lw << 16 : lw;
    }
    Vhi[t] = make_uint4(h[0], h[1], h[2], h[3]);
    Vlo[t] = make_uint4(l[0], l[1], l[2], l[3]);
}

// Rows per wave: 2 groups of 16 (every column fragment read from LDS serves
// both); 4 waves per workgroup.
constexpr int kAdjRG = 2;
constexpr int kAdjRows = 4 * 16 * kAdjRG;  // rows per workgroup
// Fast tiles: every sigma of the tile's rows and columns is 0 or at least
// kAdjFastSigma, so delta = eps/(s_i s_j) <= 1e-3; there the weight w/(1 + delta)
// is taken as w e^-delta (relative error delta^2/2 <= 5e-7) and the correction
// factor delta/(1 + delta) as delta (error delta^2 <= 1e-6 relative to q_i);
// other tiles divide.
constexpr float kAdjFastSigma = 0.01f;
constexpr float kLn2f = 0.693147180559945309f;
constexpr float kLog2ef = 1.442695040888963407f;

// Scale exponent of the dot tiles: 40, lowered for max|a| >= 2^50 so that
// 2^E Q d stays below 2^91 (sums of 64 such in FP32); 0 for inf or NaN.
__device__ __forceinline__ int adj_dot_exponent(unsigned int amax_bits)
{
    if (amax_bits == 0u) return 40;
    if (amax_bits >= 0x7f800000u) return 0;
    int ex;
    (void)frexpf(__uint_as_float(amax_bits), &ex);  // max|a| < 2^ex
    return min(max(90 - ex, 0), 40);
}

// small[g] = 1 if a block of [32g, 32g + 32) has 0 < sigma < kAdjFastSigma
// (info.x = sigma).
__global__ __launch_bounds__(256) void k_adjoint_small(const float4 *__restrict__ info, int ngroups,
                                                       int *__restrict__ small)
{
    const int g = (blockIdx.x * blockDim.x + threadIdx.x) / 32;
    if (g >= ngroups) return;
    const float s = info[g * 32 + (threadIdx.x & 31)].x;
    const unsigned long long b = __ballot(s > 0.0f && s < kAdjFastSigma);
    const unsigned int half = (threadIdx.x & 32) ? (unsigned int)(b >> 32) : (unsigned int)b;
    if ((threadIdx.x & 31) == 0) small[g] = half != 0u;
}

// s_waitcnt vmcnt(0) expcnt(0) lgkmcnt(0): this wave's global->LDS copies and
// LDS accesses are complete (issued before every barrier of the tile loop).
__device__ __forceinline__ void adj_wait_all() { __builtin_amdgcn_s_waitcnt(0); }

// The pair pass.  With u the Pearson-scaled blocks (u_i . u_j = cov/(s_i s_j)),
// d = u_i . u_j, delta = eps/(s_i s_j), w = exp2(-(t_i - t_j)^2), w~ = w/(1 + delta)
// and Q_ij = (a_i + a_j) [d > 0] w~ (symmetric in i, j), row i gathers
//   G_i   = sum_j Q_ij u_j                          (second product, MFMA)
//   q_i   = sum_j Q_ij d+ / (1 + delta)             (= sum_j Q_ij r_ij; from
//           u_i . G_i less the per-pair correction sum_j Q_ij d+ delta/(1 + delta))
//   e_i   = sum_j Q_ij d+ (t_i - t_j) + F_i
// and its block gradient is, with dL/ds_i = -q_i / s_i + cE e_i
// (cE = -2/(H tscale)),
//   (G_i / s_i + dL/ds_i u_i) / sqrt(K-1), centred over the K pixels.
// F_i = sum_j (b_i + b_j) w_ij (t_i - t_j), the part of e_i that depends on the
// sigmas only, comes from the binned Gauss expansion (adjoint_weight_terms_dev,
// FP64) and is added here per row, so the pair loop carries no weight-sum term.
//
// The u_j are centred (sum_k u_jk = 0), so G_i is too: the second product
// runs over the k-groups kg < KG = ceil((K - 1) / 16) only, and when K - 1 is a
// multiple of 16 the last component is G_{i,K-1} = -sum_{k < K-1} G_ik (at
// K = 49 three k-groups instead of four).  Its error is that of the split
// fragments summed over k (about sqrt(K) 2^-17 relative), still far below
// the FP32 rounding of the forward the gradient is checked against.
//
// info2 = (t, a, b, 1/s) per block (1/s = 0 for flat blocks and padding);
// F[v] for v < Nv.
// Up to K = 64 two workgroups share a CU; beyond, the row fragments and
// accumulators of three or four k-steps need more registers than half a
// SIMD's file (K = 128: 420 VGPRs, 130 KiB of LDS).
template <int KS, int KG>
__global__ __launch_bounds__(256, KS <= 2 ? 2 : 1) void k_adjoint_mfma(const uint4 *__restrict__ Uhi, const uint4 *__restrict__ Ulo,
                                                         const uint4 *__restrict__ Vhi, const uint4 *__restrict__ Vlo,
                                                         const float4 *__restrict__ info2, const int *__restrict__ small,
                                                         const double *__restrict__ F,
                                                         const unsigned int *__restrict__ amax_bits,
                                                         const float *__restrict__ UT,
                                                         int K, int Nv, int Nv_pad, float eps, float cE,
                                                         float inv_sqrt_km1, int K16, float *__restrict__ Gc)
{
    constexpr int RG = kAdjRG;
    constexpr int KGA = 2 * KS;                            // k-groups of a block (all K values)
    constexpr int URUN = 4 * KS * 64, VRUN = 2 * KG * 64;  // uint4 of one 64-column tile
    constexpr int NRUN = 8 * KS + 4 * KG;                  // runs of 64 uint4 copied per tile
    static_assert(NRUN % 4 == 0, "tile runs split over 4 waves");
    // two tile buffers (distinct arrays: reading one does not wait for the copy into the other)
    // two tile buffers (distinct arrays: reading one does not wait for the
    // copy into the other) up to K = 128; one beyond (K = 160: 81 KiB)
    constexpr bool DB = KS <= 4;
    __shared__ uint4 uh0[URUN], ul0[URUN], vh0[VRUN], vl0[VRUN];
    __shared__ uint4 uh1[DB ? URUN : 1], ul1[DB ? URUN : 1], vh1[DB ? VRUN : 1], vl1[DB ? VRUN : 1];
    __shared__ float4 ci0[64], ci1[DB ? 64 : 1];
    const int tid = threadIdx.x, lane = tid & 63;
    const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);
    const int li = lane & 15, lq = lane >> 4;
    const int rowbase = blockIdx.x * kAdjRows + wave * 16 * RG;

    // The row fragments carry a factor 2^E, so the dot tiles hold 2^E d and
    // [d > 0] w~ is one v_med3(2^E d, 0, w~) (0 <= w~ <= 1): the two agree
    // unless 0 < d < 2^-E, far inside the rounding of d (E = 40 but for loss
    // coefficients beyond 2^50; a pair that close to the kink of max(0, r) may
    // take either side).  corr and se then carry 2^E as well, removed per row
    // at the end.
    const int E = adj_dot_exponent(*amax_bits);
    const float dscale = __builtin_amdgcn_ldexpf(1.0f, E);
    bf16x8 bh[RG][KS], bl[RG][KS];
    float4 R[RG];
    float rinv[RG];
#pragma unroll
    for (int rg = 0; rg < RG; rg++) {
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const size_t o = ((size_t)(rowbase / 16 + rg) * KS + ks) * 64 + lane;
            bh[rg][ks] = bf16x8_scaled(__builtin_bit_cast(bf16x8, Uhi[o]), dscale);
            bl[rg][ks] = bf16x8_scaled(__builtin_bit_cast(bf16x8, Ulo[o]), dscale);
        }
        R[rg] = info2[rowbase + rg * 16 + li];
        rinv[rg] = R[rg].w;
        R[rg].w *= eps * kLog2ef;  // delta log2(e) = (eps log2(e) / s_i) (1 / s_j)
    }
    const bool row_small = small[rowbase / 32] != 0;
    const Bf16Neg negu = bf16_neg_units();
    floatx4 acc[KG][RG];
#pragma unroll
    for (int kg = 0; kg < KG; kg++)
#pragma unroll
        for (int rg = 0; rg < RG; rg++) acc[kg][rg] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
    // FP32 totals of the second product, kept in this wave's rows of Gc
    // (K16 floats per row; only this wave touches them): acc is added every
    // 128 tiles (8192 columns; the row's load waits for the tile copy in
    // flight, so the adds are kept rare).  Lane l holds row i = l % 16 of
    // each group, k = 16kg + 4(l/16) + r.
    auto flush = [&](bool first) {
#pragma unroll
        for (int rg = 0; rg < RG; rg++) {
            float *row = Gc + (size_t)(rowbase + rg * 16 + li) * K16;
#pragma unroll
            for (int kg = 0; kg < KG; kg++) {
                const int k = 16 * kg + 4 * lq;
                float4 t = first ? make_float4(0.0f, 0.0f, 0.0f, 0.0f) : *reinterpret_cast<const float4 *>(row + k);
                t.x += acc[kg][rg][0];
                t.y += acc[kg][rg][1];
                t.z += acc[kg][rg][2];
                t.w += acc[kg][rg][3];
                *reinterpret_cast<float4 *>(row + k) = t;
                acc[kg][rg] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
            }
        }
    };
    double gq[RG] = {0.0, 0.0}, ge[RG] = {0.0, 0.0};
    const int ntiles = Nv_pad / 64;

    // asynchronous copy of tile T: the column blocks' fragments hi/lo (4 KS
    // runs of 64 uint4 each), their transposed fragments hi/lo (2 KG runs
    // each) and the info run
    auto copy = [&](int T, uint4 *duh, uint4 *dul, uint4 *dvh, uint4 *dvl, float4 *dci) {
        typedef __attribute__((address_space(3))) void *lds_ptr;
        const int c0 = T * 64;
        const size_t ou = (size_t)(c0 / 16) * KS * 64, ov = (size_t)(c0 / 32) * KG * 64;
#pragma unroll
        for (int i = 0; i < NRUN / 4; i++) {
            const int run = wave + 4 * i;
            const uint4 *src;
            uint4 *dst;
            if (run < 4 * KS) { src = Uhi + ou + run * 64; dst = duh + run * 64; }
            else if (run < 8 * KS) { src = Ulo + ou + (run - 4 * KS) * 64; dst = dul + (run - 4 * KS) * 64; }
            else if (run < 8 * KS + 2 * KG) { src = Vhi + ov + (run - 8 * KS) * 64; dst = dvh + (run - 8 * KS) * 64; }
            else { src = Vlo + ov + (run - 8 * KS - 2 * KG) * 64; dst = dvl + (run - 8 * KS - 2 * KG) * 64; }
            __builtin_amdgcn_global_load_lds((const void *)(src + lane), (lds_ptr)dst, 16, 0, 0);
        }
        if (wave == 3) __builtin_amdgcn_global_load_lds((const void *)(info2 + c0 + lane), (lds_ptr)dci, 16, 0, 0);
    };

    // first-product dot tiles of column group cg for both row groups
    auto dots = [&](const uint4 *uh, const uint4 *ul, int cg, floatx4 (&d)[RG]) {
#pragma unroll
        for (int rg = 0; rg < RG; rg++) d[rg] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
        for (int ks = 0; ks < KS; ks++) {
            const bf16x8 ah = __builtin_bit_cast(bf16x8, uh[(cg * KS + ks) * 64 + lane]);
            const bf16x8 al = __builtin_bit_cast(bf16x8, ul[(cg * KS + ks) * 64 + lane]);
#pragma unroll
            for (int rg = 0; rg < RG; rg++) {
                d[rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh[rg][ks], d[rg], 0, 0, 0);
                d[rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl[rg][ks], d[rg], 0, 0, 0);
                d[rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh[rg][ks], d[rg], 0, 0, 0);
            }
        }
    };
    // pair terms of one pair (column constants C, row group rg, r of the
    // lane's four columns) into pv[h * 4 + r]:
    //   Q; corr += Q d+ (1 - 1/(1 + delta)) log2(e); se += Q d+ (t_i - t_j)
    // Fast tiles take w~ = exp2(-(t_i - t_j)^2 - delta log2(e)) = w e^-delta and
    // 1 - 1/(1 + delta) = delta (relative errors delta^2/2 and delta, delta <= 1e-3);
    // exact tiles divide.  corr is kept in log2(e) units (scaled by ln 2 per row).
    auto pair1 = [&](auto exact_tag, const float4 &C, int h, int rg, int r, float dv, float (&pv)[8], float &corr,
                     float &se) {
        constexpr bool EXACT = decltype(exact_tag)::value;
        const float dt = R[rg].x - C.x;
        const float dl = R[rg].w * C.w;
        float wt, om;
        if (EXACT) {
            const float w = __builtin_amdgcn_exp2f(-(dt * dt));
            const float kap = __builtin_amdgcn_rcpf(fmaf(dl, kLn2f, 1.0f));
            wt = w * kap;
            om = dl * kap;
        } else {
            wt = __builtin_amdgcn_exp2f(-fmaf(dt, dt, dl));
            om = dl;
        }
        // [d > 0] w~ as one v_med3 (dv = 2^E d, 0 <= w~ <= 1)
        const float Q = (R[rg].y + C.y) * __builtin_amdgcn_fmed3f(dv, 0.0f, wt);
        const float Qd = Q * dv;  // 2^E Q max(d, 0)
        corr = fmaf(Qd, om, corr);
        se = fmaf(Qd, dt, se);
        pv[h * 4 + r] = Q;
    };
    // second-product MFMAs of step st for row group rg (coefficients pv)
    auto grads = [&](const uint4 *vh, const uint4 *vl, int st, int rg, const float (&pv)[8]) {
        unsigned int ph[4], pl[4];
#pragma unroll
        for (int q = 0; q < 4; q++) split_bf16x2_dot(pv[2 * q], pv[2 * q + 1], negu, ph[q], pl[q]);
        const bf16x8 bH = __builtin_bit_cast(bf16x8, make_uint4(ph[0], ph[1], ph[2], ph[3]));
        const bf16x8 bL = __builtin_bit_cast(bf16x8, make_uint4(pl[0], pl[1], pl[2], pl[3]));
#pragma unroll
        for (int kg = 0; kg < KG; kg++) {
            const bf16x8 ah = __builtin_bit_cast(bf16x8, vh[(st * KG + kg) * 64 + lane]);
            const bf16x8 al = __builtin_bit_cast(bf16x8, vl[(st * KG + kg) * 64 + lane]);
            acc[kg][rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bH, acc[kg][rg], 0, 0, 0);
            acc[kg][rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bL, acc[kg][rg], 0, 0, 0);
            acc[kg][rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bH, acc[kg][rg], 0, 0, 0);
        }
    };
    // one 64-column tile as two 32-column steps, software-pipelined so that
    // the matrix core has independent work beside every stretch of pair terms:
    //   dots(step 0) | pairs(step 0) with dots(step 1) | pairs(step 1) with
    //   grads(step 0) | grads(step 1)
    // Inside the two middle phases the MFMAs are spread evenly between the 16
    // pairs of the lane, with scheduling fences (which only LDS reads may
    // cross) after each pair: the compiler would otherwise issue all MFMAs of
    // a phase first.
    auto tile64 = [&](auto exact_tag, const uint4 *uh, const uint4 *ul, const uint4 *vh, const uint4 *vl,
                      const float4 *ci, float (&corr)[RG], float (&se)[RG]) {
        constexpr int MDC = 3 * KS * RG;  // MFMAs of the dot tiles of one column group
        constexpr int MD = 2 * MDC;       // ... of one step
        constexpr int MG = 3 * KG * RG;   // MFMAs of the gradient product of one step
        float pv0[RG][8], pv1[RG][8];
        floatx4 d0[RG], d1[RG], dn[2][RG];
        dots(uh, ul, 0, d0);
        dots(uh, ul, 1, d1);
        __builtin_amdgcn_sched_barrier(0);
        // phase 1: pairs of step 0, dots of step 1 (column groups 2, 3)
        bf16x8 fh[2][KS], fl[2][KS];
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int ks = 0; ks < KS; ks++) {
                fh[c][ks] = __builtin_bit_cast(bf16x8, uh[((2 + c) * KS + ks) * 64 + lane]);
                fl[c][ks] = __builtin_bit_cast(bf16x8, ul[((2 + c) * KS + ks) * 64 + lane]);
            }
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int rg = 0; rg < RG; rg++) dn[c][rg] = floatx4{0.0f, 0.0f, 0.0f, 0.0f};
        auto dot_mfma = [&](int m) {
            const int c = m / MDC, ks = (m % MDC) / (3 * RG), rg = (m % (3 * RG)) / 3, sp = m % 3;
            dn[c][rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sp == 2 ? fl[c][ks] : fh[c][ks],
                                                                sp == 1 ? bl[rg][ks] : bh[rg][ks], dn[c][rg], 0, 0, 0);
        };
        float4 C[2][4];  // column constants of the step's 8 columns of this lane
#pragma unroll
        for (int c = 0; c < 2; c++)
#pragma unroll
            for (int r = 0; r < 4; r++) C[c][r] = ci[c * 16 + 4 * lq + r];
#pragma unroll
        for (int q = 0; q < 8 * RG; q++) {  // q = (rg, column group, r)
            const int rg = q / 8, cg = (q / 4) % 2, r = q % 4;
#pragma unroll
            for (int m = MD * q / (8 * RG); m < MD * (q + 1) / (8 * RG); m++) dot_mfma(m);
            pair1(exact_tag, C[cg][r], cg, rg, r, cg == 0 ? d0[rg][r] : d1[rg][r], pv0[rg], corr[rg], se[rg]);
            __builtin_amdgcn_sched_barrier(0x100);
        }
        // phase 2: pairs of step 1, gradient product of step 0
        bf16x8 gH[RG], gL[RG], vfh[KG], vfl[KG];
        {
#pragma unroll
            for (int rg = 0; rg < RG; rg++) {
                unsigned int ph[4], pl[4];
#pragma unroll
                for (int q = 0; q < 4; q++) split_bf16x2_dot(pv0[rg][2 * q], pv0[rg][2 * q + 1], negu, ph[q], pl[q]);
                gH[rg] = __builtin_bit_cast(bf16x8, make_uint4(ph[0], ph[1], ph[2], ph[3]));
                gL[rg] = __builtin_bit_cast(bf16x8, make_uint4(pl[0], pl[1], pl[2], pl[3]));
            }
#pragma unroll
            for (int kg = 0; kg < KG; kg++) {
                vfh[kg] = __builtin_bit_cast(bf16x8, vh[kg * 64 + lane]);
                vfl[kg] = __builtin_bit_cast(bf16x8, vl[kg * 64 + lane]);
            }
#pragma unroll
            for (int c = 0; c < 2; c++)
#pragma unroll
                for (int r = 0; r < 4; r++) C[c][r] = ci[(2 + c) * 16 + 4 * lq + r];
        }
        __builtin_amdgcn_sched_barrier(0x100);
        auto grad_mfma = [&](int m) {
            const int rg = m / (3 * KG), kg = (m % (3 * KG)) / 3, sp = m % 3;
            acc[kg][rg] = __builtin_amdgcn_mfma_f32_16x16x32_bf16(sp == 2 ? vfl[kg] : vfh[kg],
                                                                  sp == 1 ? gL[rg] : gH[rg], acc[kg][rg], 0, 0, 0);
        };
#pragma unroll
        for (int q = 0; q < 8 * RG; q++) {
            const int rg = q / 8, cg = (q / 4) % 2, r = q % 4;
#pragma unroll
            for (int m = MG * q / (8 * RG); m < MG * (q + 1) / (8 * RG); m++) grad_mfma(m);
            pair1(exact_tag, C[cg][r], cg, rg, r, dn[cg][rg][r], pv1[rg], corr[rg], se[rg]);
            __builtin_amdgcn_sched_barrier(0x100);
        }
        __builtin_amdgcn_sched_barrier(0);
        grads(vh, vl, 1, 0, pv1[0]);
        grads(vh, vl, 1, 1, pv1[1]);
    };

    auto compute = [&](int T, const uint4 *uh, const uint4 *ul, const uint4 *vh, const uint4 *vl, const float4 *ci) {
        const int c0 = T * 64;
        const bool exact = row_small || (small[c0 / 32] | small[c0 / 32 + 1]) != 0;
        float corr[RG], se[RG];
#pragma unroll
        for (int rg = 0; rg < RG; rg++) { corr[rg] = 0.0f; se[rg] = 0.0f; }
        if (exact) tile64(std::true_type{}, uh, ul, vh, vl, ci, corr, se);
        else tile64(std::false_type{}, uh, ul, vh, vl, ci, corr, se);
#pragma unroll
        for (int rg = 0; rg < RG; rg++) {
            gq[rg] += (double)corr[rg];
            ge[rg] += (double)se[rg];
        }
        if ((T & 127) == 127 || T == ntiles - 1) flush(T < 128);
    };

    // step T: tile T has landed (own copies waited, then the barrier), every
    // wave is done with tile T-1, whose buffer now receives tile T+1
    auto step = [&](int T, const uint4 *uh, const uint4 *ul, const uint4 *vh, const uint4 *vl, const float4 *ci,
                    uint4 *nuh, uint4 *nul, uint4 *nvh, uint4 *nvl, float4 *nci) {
        adj_wait_all();
        __syncthreads();
        if (T + 1 < ntiles) copy(T + 1, nuh, nul, nvh, nvl, nci);
        compute(T, uh, ul, vh, vl, ci);
    };
    if constexpr (DB) {
        copy(0, uh0, ul0, vh0, vl0, ci0);
        for (int T = 0; T < ntiles; T += 2) {
            step(T, uh0, ul0, vh0, vl0, ci0, uh1, ul1, vh1, vl1, ci1);
            if (T + 1 < ntiles) step(T + 1, uh1, ul1, vh1, vl1, ci1, uh0, ul0, vh0, vl0, ci0);
        }
    } else {
        for (int T = 0; T < ntiles; T++) {
            adj_wait_all();
            __syncthreads();  // every wave is done with tile T - 1
            copy(T, uh0, ul0, vh0, vl0, ci0);
            adj_wait_all();
            __syncthreads();  // tile T has landed
            compute(T, uh0, ul0, vh0, vl0, ci0);
        }
    }

    // per row i: G_k = (tot_k / s_i + dL/ds_i u_ik) / sqrt(K-1), centred over k < K.
    // Lane l holds i = l % 16 of each row group and k = 16kg + 4(l/16) + r.
    // The second product wrote tot_k for k < 16 KG (at least k < K - 1);
    // tot_{K-1} beyond that is minus the sum of the others.
    const bool derive_last = (K - 1) == 16 * KG;
#pragma unroll
    for (int rg = 0; rg < RG; rg++) {
        const int v = rowbase + rg * 16 + li;
        double q = gq[rg], e = ge[rg];
        q += __shfl_xor(q, 16);
        q += __shfl_xor(q, 32);
        e += __shfl_xor(e, 16);
        e += __shfl_xor(e, 32);
        const double unscale = __builtin_amdgcn_ldexp(1.0, -E);
        q *= (double)kLn2f * unscale;  // the correction, from log2(e) units
        e *= unscale;
        if (v < Nv) e += F[v];
        const double is = (double)rinv[rg];
        const bool live = v < Nv && rinv[rg] != 0.0f;
        float *row = Gc + (size_t)v * K16;
        double uk[KGA][4], gt[KGA][4], ug = 0.0, gsum = 0.0;
#pragma unroll
        for (int kg = 0; kg < KGA; kg++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int k = 16 * kg + 4 * lq + r;
                uk[kg][r] = (k < K) ? (double)UT[(size_t)k * Nv_pad + v] : 0.0;
                gt[kg][r] = (k < K && kg < KG) ? (double)row[k] : 0.0;
                gsum += gt[kg][r];
            }
        if (derive_last) {
            gsum += __shfl_xor(gsum, 16);
            gsum += __shfl_xor(gsum, 32);
#pragma unroll
            for (int kg = 0; kg < KGA; kg++)
#pragma unroll
                for (int r = 0; r < 4; r++)
                    if (16 * kg + 4 * lq + r == K - 1) gt[kg][r] = -gsum;
        }
#pragma unroll
        for (int kg = 0; kg < KGA; kg++)
#pragma unroll
            for (int r = 0; r < 4; r++) ug += uk[kg][r] * gt[kg][r];
        ug += __shfl_xor(ug, 16);
        ug += __shfl_xor(ug, 32);
        const double ds = -(ug - q) * is + (double)cE * e;
        double gk[KGA][4], part = 0.0;
#pragma unroll
        for (int kg = 0; kg < KGA; kg++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int k = 16 * kg + 4 * lq + r;
                gk[kg][r] = (k < K) ? (gt[kg][r] * is + ds * uk[kg][r]) * (double)inv_sqrt_km1 : 0.0;
                part += gk[kg][r];
            }
        part += __shfl_xor(part, 16);
        part += __shfl_xor(part, 32);
        const double mean = part / (double)K;
#pragma unroll
        for (int kg = 0; kg < KGA; kg++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int k = 16 * kg + 4 * lq + r;
                if (k < K16) row[k] = (live && k < K) ? (float)(gk[kg][r] - mean) : 0.0f;
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

template <int KS, int KG>
static void launch_adjoint_mfma_t(dim3 grid, hipStream_t st, const uint4 *uhi, const uint4 *ulo, const uint4 *vhi,
                                  const uint4 *vlo, const float4 *info2, const int *small, const double *F,
                                  const unsigned int *amax, const float *UT, int K, int Nv, int Nv_pad, float eps,
                                  float cE, float isk, int K16, float *Gc)
{
    hipLaunchKernelGGL((k_adjoint_mfma<KS, KG>), grid, dim3(256), 0, st, uhi, ulo, vhi, vlo, info2, small, F, amax,
                       UT, K, Nv, Nv_pad, eps, cE, isk, K16, Gc);
}

// k-groups KG = ceil((K - 1) / 16) of the second product: 1..2 (KS = 1), 2..4 (KS = 2), 4..6 (KS = 3),
// 6..8 (KS = 4)
static int launch_adjoint_mfma(int KS, int KG, dim3 grid, hipStream_t st, const uint4 *uhi, const uint4 *ulo,
                               const uint4 *vhi, const uint4 *vlo, const float4 *info2, const int *small,
                               const double *F, const unsigned int *amax, const float *UT, int K, int Nv, int Nv_pad,
                               float eps, float cE, float isk, int K16, float *Gc)
{
#define SRED_ADJ_CASE(ks, kg)                                                                                 \
    if (KS == ks && KG == kg) {                                                                               \
        launch_adjoint_mfma_t<ks, kg>(grid, st, uhi, ulo, vhi, vlo, info2, small, F, amax, UT, K, Nv, Nv_pad, \
                                      eps, cE, isk, K16, Gc);                                                 \
        return SRED_OK;                                                                                       \
    }
    SRED_ADJ_CASE(1, 1)
    SRED_ADJ_CASE(1, 2)
    SRED_ADJ_CASE(2, 2)
    SRED_ADJ_CASE(2, 3)
    SRED_ADJ_CASE(2, 4)
    SRED_ADJ_CASE(3, 4)
    SRED_ADJ_CASE(3, 5)
    SRED_ADJ_CASE(3, 6)
    SRED_ADJ_CASE(4, 6)
    SRED_ADJ_CASE(4, 7)
    SRED_ADJ_CASE(4, 8)
    SRED_ADJ_CASE(5, 8)
    SRED_ADJ_CASE(5, 9)
    SRED_ADJ_CASE(5, 10)
#undef SRED_ADJ_CASE
    set_error("internal: no adjoint pair kernel for this block size");
    return SRED_ERR_ARG;
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
    // forward: statistics, noise variance, mask, accumulators and the normalised map
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
    p.precision = SRED_BF16X3;  // the forward map: split-BF16 symmetric contraction
    p.filter_param = nv;
    p.threshold = 0.0f;
    SRED_TRY(global_accum_dev(ctx, g, p, d_img, d_means, d_stds, d_S, d_W));
    SRED_TRY(global_finalize_dev(ctx, g, SRED_PEARSON, d_S, d_W, d_mask, np, 0.0f, 1, d_map));
    const unsigned int *d_mm = reinterpret_cast<const unsigned int *>(ctx->misc.as<char>() + 1024);
    const int Nv = ctx->last_nv;
    SRED_TRY(zero_dev(ctx, d_grad, sizeof(float) * n));
    if (Nv == 0) return SRED_OK;
    const int KS = (K + 31) / 32;
    const int Nv_pad = (Nv + kAdjRows - 1) / kAdjRows * kAdjRows;
    // blocks in ascending sigma: the few tiles that need the exact small-sigma
    // form then hold the small-sigma blocks only (any order gives the same
    // gradient up to rounding; pix2v below follows it)
    const int *d_idx = ctx->valid_idx.as<int>();
    if (Nv > 1) {
        int *sorted = nullptr;
        SRED_TRY(sort_valid_by_sigma(ctx, ctx->sortbuf, d_stds, d_idx, Nv, &sorted));
        d_idx = sorted;
    }

    const double H = java_float_literal(nv) + java_float_literal(kEps);
    const double tscale_d = std::sqrt(1.4426950408889634 / H);
    const float tscale = (float)tscale_d;
    const float inv_sqrt_km1 = (float)(1.0 / std::sqrt((double)(K - 1)));
    SRED_TRY(ctx->aux.ensure(2 * sizeof(float4) * (size_t)Nv_pad));
    float4 *d_info = ctx->aux.as<float4>(), *d_info2 = d_info + Nv_pad;
    // F (FP64 per block) and the bits of max |b| behind it
    SRED_TRY(ctx->adj_f.ensure(sizeof(double) * ((size_t)Nv_pad + 1)));
    double *d_F = ctx->adj_f.as<double>();
    unsigned int *d_bmax = reinterpret_cast<unsigned int *>(d_F + Nv_pad);
    SRED_TRY(zero_dev(ctx, d_bmax, sizeof(double)));
    hipLaunchKernelGGL(k_adjoint_coef, dim3((Nv_pad + 255) / 256), dim3(256), 0, ctx->stream, d_idx, Nv, Nv_pad,
                       d_stds, d_S, d_W, d_mask, d_dmap, d_mm, np, tscale, d_info, d_info2, d_bmax);
    // Pearson-scaled block vectors, k-major
    SRED_TRY(upload_offsets(ctx, g, g.w, &ctx->offs2, ctx->offs2_key));
    size_t nu = (size_t)Nv_pad * K4;
    SRED_TRY(ctx->u32.ensure(sizeof(float) * nu));
    launch_gather_f32(ctx, SRED_PEARSON, d_img, d_means, d_stds, d_idx, Nv, Nv_pad, K, K4, ctx->offs2.as<int>(),
                      inv_sqrt_km1, (float)(1.0 / std::sqrt((double)K)), ctx->u32.as<float>());
    SRED_TRY(ctx->partial.ensure(sizeof(float) * (size_t)Nv_pad * K16));
    float *d_Gc = ctx->partial.as<float>();
    {
        // second product over k < 16 KG (the last component follows from sum_k u_k = 0)
        const int KG = std::max(1, (K - 1 + 15) / 16);
        const size_t u4 = (size_t)Nv_pad * KS * 4;  // uint4 per fragment-major array
        const size_t v4 = (size_t)Nv_pad * KG * 2;  // uint4 per transposed array
        SRED_TRY(ctx->uhi.ensure(sizeof(uint4) * u4));
        SRED_TRY(ctx->ulo.ensure(sizeof(uint4) * u4));
        SRED_TRY(ctx->adj_v.ensure(2 * sizeof(uint4) * v4));
        // the sigma-only terms of e_i (Gauss expansion, FP64)
        SRED_TRY(adjoint_weight_terms_dev(ctx, ctx->wmom, ctx->wpart, d_stds, d_idx, d_info2, d_bmax, Nv, tscale_d,
                                          ctx->last_sigma_max, d_F));
        const int ngroups = Nv_pad / 32;
        SRED_TRY(ctx->tflags.ensure(sizeof(int) * (size_t)ngroups));
        uint4 *d_uhi = ctx->uhi.as<uint4>(), *d_ulo = ctx->ulo.as<uint4>();
        uint4 *d_vhi = ctx->adj_v.as<uint4>(), *d_vlo = d_vhi + v4;
        int *d_small = ctx->tflags.as<int>();
        launch_gather_bf16(ctx, SRED_PEARSON, d_img, d_means, d_stds, d_idx, Nv, Nv_pad, K, KS,
                           ctx->offs2.as<int>(), inv_sqrt_km1, (float)(1.0 / std::sqrt((double)K)), d_uhi, d_ulo);
        hipLaunchKernelGGL(k_transpose_frag, dim3((unsigned)((v4 + 255) / 256)), dim3(256), 0, ctx->stream, d_uhi,
                           d_ulo, Nv_pad, KS, KG, d_vhi, d_vlo);
        hipLaunchKernelGGL(k_adjoint_small, dim3((ngroups * 32 + 255) / 256), dim3(256), 0, ctx->stream, d_info,
                           ngroups, d_small);
        SRED_TRY(time_begin(ctx));
        const float cE = (float)(-2.0 / H / (double)tscale);  // (s_i - s_j) = (t_i - t_j) / tscale
        const float *UTp = ctx->u32.as<float>();
        const dim3 grid(Nv_pad / kAdjRows);
        SRED_TRY(launch_adjoint_mfma(KS, KG, grid, ctx->stream, d_uhi, d_ulo, d_vhi, d_vlo, d_info2, d_small, d_F,
                                     d_bmax + 1, UTp, K, Nv, Nv_pad, kEps, cE, inv_sqrt_km1, K16, d_Gc));
        SRED_HIP_CHECK(hipGetLastError());
        SRED_TRY(time_end(ctx, (int64_t)Nv_pad * Nv_pad));
    }
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
