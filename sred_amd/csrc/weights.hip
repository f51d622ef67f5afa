// weights.hip -- weight sums W of the global repetition maps.
//
// Every global kernel of the reference accumulates, per reference pixel,
// Wsum_i = sum over valid test blocks j of w_ij, w_ij = exp(-(s_i - s_j)^2 /
// (h + eps)) (kernelGetGlobalPearson2D.cl:92-95; the SSIM, NRMSE and
// AbsDiffStds kernels use the same weight).  It depends on the block standard
// deviations only: with t = s sqrt(log2(e) / (h + eps)), w_ij =
// exp2(-(t_i - t_j)^2), a one-dimensional Gaussian sum, so it does not have to
// ride along in the N x N contraction.  It is evaluated here in FP64 by a
// binned Hermite expansion (the fast Gauss transform):
//
//   * blocks are binned by t, bin b centred at c_b = b * kBinW; a bin keeps
//     the moments M_n = sum_j u_j^n, u_j = (t_j - c_b) / (kBinW / 2) in [-1, 1);
//   * with a = t_i - c_b, x = sqrt(ln 2) a and z = sqrt(ln 2) kBinW / 2, the
//     generating function of the Hermite polynomials, exp(2xy - y^2) =
//     sum_n H_n(x) y^n / n!, gives
//       sum_{j in b} exp2(-(a - u_j kBinW / 2)^2) = exp2(-a^2) sum_n c_n M_n,
//       c_n = H_n(x) z^n / n!,  c_{n+1} = (2 x z c_n - 2 z^2 c_{n-1}) / (n + 1);
//   * with kBinW = 1/2 and kMoments = 16 terms the truncation is below 1e-15
//     of a bin's weight (Cramer's bound on H_n); bins with |a| > kRadius hold
//     weights below 2^-49 each and are skipped (< 2e-9 relative over 10^6
//     blocks, against W_i >= 1 from the block itself).
//
// Moments are accumulated in 2^-42 fixed point with 64-bit integer atomics
// (per-workgroup LDS tables, then an ordered sum over workgroups): integer
// sums do not depend on the order of the additions, so W is deterministic.
// Beyond 2^20 blocks the fixed point steps up (2^-41 to 2^-(62 - ceil(log2 N)))
// so that a bin holding every block still sums below 2^62.
#include <cmath>

#include "sred_internal.h"

namespace sred {

constexpr int kMoments = 16;
constexpr double kBinW = 0.5;
constexpr double kRadius = 7.25;
constexpr int kChunkBins = 256;            // bins per moments pass (LDS table 32 KiB)
constexpr int kMomentGroups = 512;         // workgroups of a moments pass
constexpr int kMaxBins = 1 << 20;
constexpr double kSqrtLn2 = 0.83255461115769775635;

__global__ __launch_bounds__(256) void k_wmoments(const float *__restrict__ stds, const int *__restrict__ idx, int Nv,
                                                  double tscale, int bin0, int nb, double fix,
                                                  unsigned long long *__restrict__ part)
{
    __shared__ unsigned long long tab[kChunkBins * kMoments];
    for (int i = threadIdx.x; i < kChunkBins * kMoments; i += 256) tab[i] = 0ull;
    __syncthreads();
    for (int v = blockIdx.x * 256 + threadIdx.x; v < Nv; v += gridDim.x * 256) {
        const double t = (double)stds[idx[v]] * tscale;
        const double q = floor(t / kBinW + 0.5);
        const int b = (int)q - bin0;
        if (b < 0 || b >= nb) continue;
        const double u = (t - q * kBinW) * (2.0 / kBinW);
        double pw = 1.0;
#pragma unroll
        for (int n = 0; n < kMoments; n++) {
            atomicAdd(&tab[b * kMoments + n], (unsigned long long)llrint(pw * fix));
            pw *= u;
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kChunkBins * kMoments; i += 256)
        part[(size_t)blockIdx.x * kChunkBins * kMoments + i] = tab[i];
}

__global__ __launch_bounds__(256) void k_wmoments_sum(const unsigned long long *__restrict__ part, int ngroups,
                                                      int bin0, int nb, double fix, double *__restrict__ M)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nb * kMoments) return;
    unsigned long long s = 0ull;
    for (int g = 0; g < ngroups; g++) s += part[(size_t)g * kChunkBins * kMoments + i];
    M[(size_t)bin0 * kMoments + i] = (double)(long long)s / fix;
}

__global__ __launch_bounds__(256) void k_weval(const float *__restrict__ stds, const int *__restrict__ idx, int Nv,
                                               double tscale, const double *__restrict__ M, int nbins,
                                               float *__restrict__ W_map)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= Nv) return;
    const int p = idx[v];
    const double t = (double)stds[p] * tscale;
    const double z = kSqrtLn2 * (kBinW / 2.0);
    const int blo = max((int)floor((t - kRadius) / kBinW + 0.5), 0);
    const int bhi = min((int)floor((t + kRadius) / kBinW + 0.5), nbins - 1);
    double W = 0.0;
    for (int b = blo; b <= bhi; b++) {
        const double *m = M + (size_t)b * kMoments;
        if (m[0] == 0.0) continue;
        const double a = t - (double)b * kBinW;
        const double xz2 = 2.0 * kSqrtLn2 * a * z, zz2 = 2.0 * z * z;
        double cm1 = 1.0, c = xz2;
        double acc = fma(c, m[1], m[0]);
#pragma unroll
        for (int n = 1; n + 1 < kMoments; n++) {
            const double cn1 = (xz2 * c - zz2 * cm1) / (double)(n + 1);
            acc = fma(cn1, m[n + 1], acc);
            cm1 = c;
            c = cn1;
        }
        W = fma(exp2(-a * a), acc, W);
    }
    W_map[p] = (float)W;
}

// Fixed-point scale of the moment sums: 2^42 up to 2^20 blocks, then
// 2^(62 - ceil(log2 Nv)), so that Nv terms of magnitude <= 1 sum below 2^62.
static double fixed_point_scale(int Nv)
{
    int L = 20;
    while (L < 31 && (1LL << L) < (long long)Nv) L++;
    return std::ldexp(1.0, 62 - L);
}

// W_map[idx[v]] for the Nv valid blocks; sigma_max bounds their standard
// deviations (it sizes the bin table).
int weight_sums_dev(sred_ctx *ctx, DevBuf &wmom, DevBuf &wpart, const float *d_stds, const int *d_idx, int Nv,
                    double tscale, float sigma_max, float *d_W)
{
    if (Nv <= 0) return SRED_OK;
    const double tmax = (double)sigma_max * tscale;
    if (tmax != tmax) {
        set_error("non-finite block statistics (NaN or infinite values in the image)");
        return SRED_ERR_NONFINITE;
    }
    if (!(tmax < (double)(kMaxBins - 2) * kBinW)) {
        set_error("block standard deviations span too wide a range of weights (normalise the image)");
        return SRED_ERR_ARG;
    }
    const int nbins = (int)std::floor(tmax / kBinW + 0.5) + 1;
    const double fix = fixed_point_scale(Nv);
    SRED_TRY(wmom.ensure(sizeof(double) * (size_t)nbins * kMoments));
    SRED_TRY(wpart.ensure(sizeof(unsigned long long) * (size_t)kMomentGroups * kChunkBins * kMoments));
    double *d_M = wmom.as<double>();
    unsigned long long *d_part = wpart.as<unsigned long long>();
    const int groups = std::min(kMomentGroups, (Nv + 255) / 256);
    for (int bin0 = 0; bin0 < nbins; bin0 += kChunkBins) {
        const int nb = std::min(kChunkBins, nbins - bin0);
        hipLaunchKernelGGL(k_wmoments, dim3(groups), dim3(256), 0, ctx->stream, d_stds, d_idx, Nv, tscale, bin0, nb,
                           fix, d_part);
        hipLaunchKernelGGL(k_wmoments_sum, dim3((nb * kMoments + 255) / 256), dim3(256), 0, ctx->stream, d_part,
                           groups, bin0, nb, fix, d_M);
    }
    hipLaunchKernelGGL(k_weval, dim3((Nv + 255) / 256), dim3(256), 0, ctx->stream, d_stds, d_idx, Nv, tscale, d_M,
                       nbins, d_W);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

// ---------------------------------------------------------------------------
// The sigma-only term of the global Pearson adjoint.  The weight sums W_i
// enter the map through rep_i = S_i / (W_i n + eps); with b_i the cotangent of
// W_i (one value per block, any sign) the pairs contribute, to dL/dt_i,
//   F_i = sum_j (b_i + b_j) w_ij (t_i - t_j),  w_ij = exp2(-(t_i - t_j)^2),
// which again depends on the block sigmas only.  It is evaluated with the same
// binned expansion and two source weights, s_j = 1 and s_j = b_j: with
// a = t_i - c_b and h = kBinW / 2, t_i - t_j = a - u_j h, so
//   sum_{j in b} s_j w_ij (t_i - t_j) = exp2(-a^2) (a sum_n c_n M_n - h sum_n c_n M_{n+1}),
// M_n = sum_{j in b} s_j u_j^n, n = 0..kMoments (one moment more than W).
// The b moments are taken in fixed point relative to 2^e >= max |b_j|.
constexpr int kAdjMoments = kMoments + 1;  // M_0 .. M_kMoments
constexpr int kAdjChunkBins = 128;        // bins per pass (LDS table 2 x 17 x 128 x 8 B = 34 KiB)

// scale s with |b| s <= 1 for every block (0 when every b is 0; NaN for a non-finite b)
__device__ __forceinline__ double adj_b_scale(unsigned int bmax_bits)
{
    const float bmax = __uint_as_float(bmax_bits);
    if (bmax_bits >= 0x7f800000u) return __longlong_as_double(0x7ff8000000000000LL);  // inf or NaN
    return bmax > 0.0f ? exp2(-ceil(log2((double)bmax))) : 0.0;
}

__global__ __launch_bounds__(256) void k_adj_wmoments(const float *__restrict__ stds, const int *__restrict__ idx,
                                                      const float4 *__restrict__ info2, int Nv, double tscale,
                                                      const unsigned int *__restrict__ bmax_bits, double fix, int bin0,
                                                      int nb, unsigned long long *__restrict__ part)
{
    __shared__ unsigned long long tab[kAdjChunkBins * 2 * kAdjMoments];
    for (int i = threadIdx.x; i < kAdjChunkBins * 2 * kAdjMoments; i += 256) tab[i] = 0ull;
    __syncthreads();
    const double bs = adj_b_scale(*bmax_bits);
    if (bs == bs) {
        for (int v = blockIdx.x * 256 + threadIdx.x; v < Nv; v += gridDim.x * 256) {
            const double t = (double)stds[idx[v]] * tscale;
            const double q = floor(t / kBinW + 0.5);
            const int b = (int)q - bin0;
            if (b < 0 || b >= nb) continue;
            const double u = (t - q * kBinW) * (2.0 / kBinW);
            const double sb = (double)info2[v].z * bs;
            unsigned long long *t1 = tab + (size_t)(2 * b) * kAdjMoments, *tb = t1 + kAdjMoments;
            double pw = 1.0;
#pragma unroll
            for (int n = 0; n < kAdjMoments; n++) {
                atomicAdd(&t1[n], (unsigned long long)llrint(pw * fix));
                atomicAdd(&tb[n], (unsigned long long)llrint(sb * pw * fix));
                pw *= u;
            }
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < kAdjChunkBins * 2 * kAdjMoments; i += 256)
        part[(size_t)blockIdx.x * kAdjChunkBins * 2 * kAdjMoments + i] = tab[i];
}

// M[(bin * 2 + s) * kAdjMoments + n]: s = 0 the unit sources, s = 1 the b sources (unscaled)
__global__ __launch_bounds__(256) void k_adj_wmoments_sum(const unsigned long long *__restrict__ part, int ngroups,
                                                          int bin0, int nb, double fix,
                                                          const unsigned int *__restrict__ bmax_bits,
                                                          double *__restrict__ M)
{
    const int i = blockIdx.x * 256 + threadIdx.x;
    if (i >= nb * 2 * kAdjMoments) return;
    unsigned long long s = 0ull;
    for (int g = 0; g < ngroups; g++) s += part[(size_t)g * kAdjChunkBins * 2 * kAdjMoments + i];
    double m = (double)(long long)s / fix;
    if ((i / kAdjMoments) & 1) {
        const double bs = adj_b_scale(*bmax_bits);
        m = bs > 0.0 ? m / bs : (bs == bs ? 0.0 : bs);
    }
    M[(size_t)bin0 * 2 * kAdjMoments + i] = m;
}

// F[v] = b_v E1_v + Eb_v, E1 = sum_j w (t_v - t_j), Eb = sum_j b_j w (t_v - t_j)
__global__ __launch_bounds__(256) void k_adj_weval(const float *__restrict__ stds, const int *__restrict__ idx,
                                                   const float4 *__restrict__ info2, int Nv, double tscale,
                                                   const double *__restrict__ M, int nbins, double *__restrict__ F)
{
    const int v = blockIdx.x * 256 + threadIdx.x;
    if (v >= Nv) return;
    const double t = (double)stds[idx[v]] * tscale;
    const double z = kSqrtLn2 * (kBinW / 2.0), h = kBinW / 2.0;
    const int blo = max((int)floor((t - kRadius) / kBinW + 0.5), 0);
    const int bhi = min((int)floor((t + kRadius) / kBinW + 0.5), nbins - 1);
    double E1 = 0.0, Eb = 0.0;
    for (int b = blo; b <= bhi; b++) {
        const double *m1 = M + (size_t)(2 * b) * kAdjMoments, *mb = m1 + kAdjMoments;
        if (m1[0] == 0.0) continue;
        const double a = t - (double)b * kBinW;
        const double xz2 = 2.0 * kSqrtLn2 * a * z, zz2 = 2.0 * z * z;
        // n = 0 and 1 terms, then the recursion up to n = kMoments - 1
        double cm1 = 1.0, c = xz2;
        double s1 = fma(c, m1[1], m1[0]), s1u = fma(c, m1[2], m1[1]);
        double sb = fma(c, mb[1], mb[0]), sbu = fma(c, mb[2], mb[1]);
#pragma unroll
        for (int n = 1; n + 1 < kMoments; n++) {
            const double cn1 = (xz2 * c - zz2 * cm1) / (double)(n + 1);
            s1 = fma(cn1, m1[n + 1], s1);
            s1u = fma(cn1, m1[n + 2], s1u);
            sb = fma(cn1, mb[n + 1], sb);
            sbu = fma(cn1, mb[n + 2], sbu);
            cm1 = c;
            c = cn1;
        }
        const double e = exp2(-a * a);
        E1 = fma(e, fma(a, s1, -h * s1u), E1);
        Eb = fma(e, fma(a, sb, -h * sbu), Eb);
    }
    F[v] = fma((double)info2[v].z, E1, Eb);
}

// F[v] for the Nv valid blocks in adjoint order (info2[v] = (t, a, b, 1/s));
// *d_bmax holds the bits of max_v |b_v|; sigma_max bounds the block sigmas.
int adjoint_weight_terms_dev(sred_ctx *ctx, DevBuf &wmom, DevBuf &wpart, const float *d_stds, const int *d_idx,
                             const float4 *d_info2, const unsigned int *d_bmax, int Nv, double tscale,
                             float sigma_max, double *d_F)
{
    if (Nv <= 0) return SRED_OK;
    const double tmax = (double)sigma_max * tscale;
    if (tmax != tmax) {
        set_error("non-finite block statistics (NaN or infinite values in the image)");
        return SRED_ERR_NONFINITE;
    }
    if (!(tmax < (double)(kMaxBins - 2) * kBinW)) {
        set_error("block standard deviations span too wide a range of weights (normalise the image)");
        return SRED_ERR_ARG;
    }
    const int nbins = (int)std::floor(tmax / kBinW + 0.5) + 1;
    const double fix = fixed_point_scale(Nv);
    SRED_TRY(wmom.ensure(sizeof(double) * (size_t)nbins * 2 * kAdjMoments));
    SRED_TRY(wpart.ensure(sizeof(unsigned long long) * (size_t)kMomentGroups * kAdjChunkBins * 2 * kAdjMoments));
    double *d_M = wmom.as<double>();
    unsigned long long *d_part = wpart.as<unsigned long long>();
    const int groups = std::min(kMomentGroups, (Nv + 255) / 256);
    for (int bin0 = 0; bin0 < nbins; bin0 += kAdjChunkBins) {
        const int nb = std::min(kAdjChunkBins, nbins - bin0);
        hipLaunchKernelGGL(k_adj_wmoments, dim3(groups), dim3(256), 0, ctx->stream, d_stds, d_idx, d_info2, Nv,
                           tscale, d_bmax, fix, bin0, nb, d_part);
        hipLaunchKernelGGL(k_adj_wmoments_sum, dim3((nb * 2 * kAdjMoments + 255) / 256), dim3(256), 0, ctx->stream,
                           d_part, groups, bin0, nb, fix, d_bmax, d_M);
    }
    hipLaunchKernelGGL(k_adj_weval, dim3((Nv + 255) / 256), dim3(256), 0, ctx->stream, d_stds, d_idx, d_info2, Nv,
                       tscale, d_M, nbins, d_F);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

}  // namespace sred
