// pairmaps.hip -- all-pairs maps on the vector ALU: the global Hu map and
// the phase correlation map.
//
// Both are OpenCL kernels of the reference without a Java caller:
//   * kernelGetHuMap.cl:15-172 -- map[i] = sum over interior j of
//     |hu_j - hu_i| * w(s_i, s_j), hu = nu20 + nu02 of the mean-subtracted
//     bw x bh rectangle (getInvariant, :146-172), w = (100^sim - 1) / 99,
//     sim = 1 - |s_i - s_j| / (|s_i| + |s_j| + eps) (:136-144);
//   * kernelGetPhaseCorrelationMap.cl:13-179 -- does not compile (nUnique is
//     undefined at :27) and its DFT phases use integer division (:66-67).
//     The specification implemented here is textbook phase correlation on the
//     same blocks: b = (pixel - mean) / (std + eps) on the rectangle, unit
//     spectra U = F / (|F| + eps) of the 2D DFT F, R = U_i conj(U_j) with i
//     the lower interior index of the pair (one displacement per pair),
//     r = Re IDFT(R), the first strict positive maximum of r in raster order,
//     its coordinates wrapped to signed displacements d (a circular shift s
//     of block j against block i peaks at -s), map[i] = sum over interior j
//     of |d_ij| * w(s_i, s_j) with the kernel's weight (:167-178): 1 if
//     s_i == s_j, else 1 - 2 |s_i - s_j| / (|s_i| + |s_j|), in float.
// Both take the reference pixels from an optional list (the kernels'
// uniqueStdCoords "speedUp" list, :27-39); others stay 0.
//
// Layout.  Columns are all interior pixels in raster order (interior index
// j = (y - brh) * iw + (x - brw)); rows are the reference pixels.  One lane
// owns one row and walks a chunk of columns in order with the column data
// wave-uniform (scalar loads), so a row's sum is a sequential per-lane sum;
// chunks of columns run as a second grid dimension when rows alone do not
// fill the GPU, and their FP64 partial sums are added in chunk order.  The
// full Hu map (every interior pixel a row) evaluates each unordered pair
// once instead (k_hu_pairs_sym).
//
//   * Hu: per pair one reciprocal and one exp2 on FP32 (transcendental-rate
//     bound, ~15 VALU issue slots per pair), 64-column FP32 sub-sums added
//     into an FP64 row sum.
//   * Phase: FP64 throughout.  Per pair R on the half-plane of frequencies
//     ky <= bh/2 (NT = bw (bh/2 + 1) complex products), then r row by row:
//     V(kx, y) = R(kx, 0) + 2 sum_{ky >= 1} R(kx, ky) e^{2 pi i ky y / bh} and
//     r(x, y) = Re sum_kx V(kx, y) e^{2 pi i kx x / bw}, which is the full
//     inverse DFT because R(-k) = conj R(k) for real blocks.  Block sides are
//     template parameters (odd, 3..11), so all loops unroll and twiddle
//     indices are compile-time constants.
#include <float.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "sred_internal.h"

namespace sred {

constexpr int kPmTile = 16;
constexpr float kLog2_100 = 6.64385618977472469574f;
constexpr int kMaxPhaseSide = 11;

// ---------------------------------------------------------------------------
// Shared: row list and chunking
// ---------------------------------------------------------------------------

struct PairRows {
    int R = 0;                    // rows
    const int *d_rowpix = nullptr;  // pixel index per row (nullptr: row r is interior index r)
    int nchunks = 1, chunk = 0;
};

// Interior pixels of the optional host list (order kept), uploaded to buf.
static int setup_rows(sred_ctx *ctx, const Geometry &g, const int32_t *rows, int64_t nrows, DevBuf &buf,
                      PairRows *pr)
{
    const int iw = g.w - 2 * g.brw, ih = g.h - 2 * g.brh;
    if (!rows) {
        pr->R = iw * ih;
        pr->d_rowpix = nullptr;
        return SRED_OK;
    }
    if (nrows < 0) {
        set_error("negative row count");
        return SRED_ERR_ARG;
    }
    std::vector<int> keep;
    keep.reserve((size_t)nrows);
    for (int64_t r = 0; r < nrows; r++) {
        const int64_t p = rows[r];
        if (p < 0 || p >= g.n()) continue;
        const int x = (int)(p % g.w), y = (int)(p / g.w);
        if (x < g.brw || x >= g.w - g.brw || y < g.brh || y >= g.h - g.brh) continue;  // :41-44
        keep.push_back((int)p);
    }
    pr->R = (int)keep.size();
    if (keep.empty()) return SRED_OK;
    SRED_TRY(buf.ensure(sizeof(int) * keep.size()));
    SRED_HIP_CHECK(hipMemcpyAsync(buf.ptr, keep.data(), sizeof(int) * keep.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
    // the host list must outlive the copy
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    pr->d_rowpix = buf.as<int>();
    return SRED_OK;
}

// Column chunks so that rows x chunks gives at least `waves` waves.
static void plan_chunks(PairRows *pr, int ncols, int waves)
{
    const int row_waves = std::max(1, (pr->R + 63) / 64);
    int nch = std::max(1, (waves + row_waves - 1) / row_waves);
    nch = std::min(nch, std::max(1, ncols / 256));
    int chunk = (ncols + nch - 1) / nch;
    chunk = (chunk + 63) / 64 * 64;
    pr->chunk = std::max(chunk, 64);
    pr->nchunks = std::max(1, (ncols + pr->chunk - 1) / pr->chunk);
}

__device__ __forceinline__ int row_pixel(const int *rowpix, int r, int iw, int w, int brw, int brh)
{
    if (rowpix) return rowpix[r];
    const int y = r / iw;
    return (brh + y) * w + brw + (r - y * iw);
}

__device__ __forceinline__ int interior_index(int p, int iw, int w, int brw, int brh)
{
    const int y = p / w;
    return (y - brh) * iw + (p - y * w - brw);
}

// map[pixel of row r] = scale * sum over chunks (in order) of part[c][r].
__global__ __launch_bounds__(256) void k_pm_finish(const double *__restrict__ part, int R, int nchunks,
                                                   const int *__restrict__ rowpix, int iw, int w, int brw, int brh,
                                                   double scale, float *__restrict__ map)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= R) return;
    double s = 0.0;
    for (int c = 0; c < nchunks; c++) s += part[(size_t)c * R + r];
    map[row_pixel(rowpix, r, iw, w, brw, brh)] = (float)(s * scale);
}

// ---------------------------------------------------------------------------
// Global Hu map
// ---------------------------------------------------------------------------

// hu = inv20 + inv02 of getInvariant (kernelGetHuMap.cl:146-172) at every
// interior pixel, in the kernel's float order: patch = pixel - mean
// (:62-72), mu_pq += (dx^p dy^q) * patch with dx = i + 1 - x_avg (an
// integer, so the powers are exact), mu_00 += patch; mu_00^2 rounded to
// float, "+ EPSILON" in double (an unsuffixed macro literal), quotient
// rounded to float.  Writes the map (border 0) and/or the column record
// (hu, std + EPSILON/2) of every interior pixel (either pointer may be null):
// with s' = s + EPSILON/2 the pair weight's denominator |s_i| + |s_j| + EPSILON
// (:141, stds are non-negative) is s'_i + s'_j, one addition per pair fewer
// and still symmetric in (i, j) bit for bit; the rounding of s' moves
// |s_i - s_j| by at most one ulp of the larger s (a relative weight change
// below 2e-7 at s ~ 0.05).
__global__ __launch_bounds__(256) void k_hu_invariants(const float *__restrict__ img, const float *__restrict__ means,
                                                       const float *__restrict__ stds, int w, int h, int bw, int bh,
                                                       float xavg, float yavg, double eps,
                                                       float *__restrict__ hu_map, float2 *__restrict__ cols)
{
#pragma clang fp contract(off)
    extern __shared__ float tile[];
    const int brw = bw / 2, brh = bh / 2;
    const int tw = kPmTile + 2 * brw, th = kPmTile + 2 * brh;
    const int x0 = blockIdx.x * kPmTile, y0 = blockIdx.y * kPmTile;
    for (int i = threadIdx.x; i < tw * th; i += blockDim.x) {
        const int ly = i / tw, lx = i - ly * tw;
        const int gy = y0 - brh + ly, gx = x0 - brw + lx;
        tile[i] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? img[(int64_t)gy * w + gx] : 0.0f;
    }
    __syncthreads();
    const int tx = threadIdx.x % kPmTile, ty = threadIdx.x / kPmTile;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx >= w || gy >= h) return;
    const int64_t p = (int64_t)gy * w + gx;
    if (gx < brw || gx >= w - brw || gy < brh || gy >= h - brh) {
        if (hu_map) hu_map[p] = 0.0f;
        return;
    }
    const float mean = means[p];
    const float *c = tile + ty * tw + tx;
    float mu20 = 0.0f, mu02 = 0.0f, mu00 = 0.0f;
    for (int j = 0; j < bh; j++) {
        const float dy = (float)j + 1.0f - yavg;
        const float dy2 = 1.0f * (dy * dy);
        for (int i = 0; i < bw; i++) {
            const float v = c[j * tw + i] - mean;
            const float dx = (float)i + 1.0f - xavg;
            mu20 += ((dx * dx) * 1.0f) * v;
            mu02 += dy2 * v;
            mu00 += v;
        }
    }
    const float m2 = mu00 * mu00;
    const float inv20 = (float)((double)mu20 / ((double)m2 + eps));
    const float inv02 = (float)((double)mu02 / ((double)m2 + eps));
    const float hu = inv20 + inv02;
    if (hu_map) hu_map[p] = hu;
    if (cols) cols[(int64_t)(gy - brh) * (w - 2 * brw) + (gx - brw)] = make_float2(hu, stds[p] + 0.5f * kEps);
}

// Row sums of |hu_j - hu_i| (100^sim - 1) over the column chunk blockIdx.y
// (the 1/99 is applied by k_pm_finish).  The column records are read with
// wave-uniform addresses.
__global__ __launch_bounds__(256) void k_hu_pairs(const float2 *__restrict__ cols, int ncols,
                                                  const int *__restrict__ rowpix, int R, int iw, int w, int brw,
                                                  int brh, int chunk, double *__restrict__ part)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= R) return;
    const float2 me = cols[interior_index(row_pixel(rowpix, r, iw, w, brw, brh), iw, w, brw, brh)];
    const float hi = me.x, si = me.y;
    const int c0 = blockIdx.y * chunk, c1 = min(c0 + chunk, ncols);
    double acc = 0.0;
    for (int j0 = c0; j0 < c1; j0 += 64) {
        const int jn = min(j0 + 64, c1);
        float a = 0.0f;
#pragma unroll 8
        for (int j = j0; j < jn; j++) {
            const float2 cj = cols[j];
            const float q = fabsf(si - cj.y) * __builtin_amdgcn_rcpf(si + cj.y);
            const float pw = __builtin_amdgcn_exp2f(fmaf(-kLog2_100, q, kLog2_100));  // 100^(1 - q)
            a = fmaf(fabsf(cj.x - hi), pw - 1.0f, a);
        }
        acc += (double)a;
    }
    part[(size_t)blockIdx.y * R + r] = acc;
}

// Symmetric form of k_hu_pairs for the full map (every interior pixel a
// row).  The pair term is symmetric in (i, j) bit for bit (|hu_j - hu_i|,
// |s_i - s_j| and s_i + s_j are), so each unordered pair of panels (p, c),
// c >= p, of kHuPanel interior pixels is evaluated once: one workgroup per
// panel pair, thread t holding rows p P + t + 256 k (k < kHuRows).  Per
// column j of panel c (wave-uniform) a thread adds its rows' terms to their
// FP32 sub-sums (into FP64 every 64 columns) and, off the diagonal, to one
// column sum, which the wave reduces (butterfly) into its own LDS row of
// column sums.  Row totals go to part[c][i], column totals (the four waves'
// sums in FP64) to part[p][j]: one writer per slot, summed in slot order by
// k_pm_finish.  Diagonal panels evaluate their ordered pairs for the rows
// only.  FULL: every row of the panel is an interior pixel (all but the
// last row panel).
constexpr int kHuPanel = 2048, kHuRows = kHuPanel / 256;
template <bool FULL>
__device__ __forceinline__ void hu_sym_panel(const float2 *__restrict__ cols, int N, int p, int c,
                                             double *__restrict__ part, float (*colacc)[kHuPanel])
{
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const bool diag = p == c;
    const int row0 = p * kHuPanel, col0 = c * kHuPanel, ncol = min(kHuPanel, N - col0);
    float hi[kHuRows], si[kHuRows], r32[kHuRows];
    double r64[kHuRows];
#pragma unroll
    for (int k = 0; k < kHuRows; k++) {
        const int r = row0 + tid + 256 * k;
        const float2 me = (FULL || r < N) ? cols[r] : make_float2(0.0f, 0.0f);
        hi[k] = me.x;
        si[k] = me.y;
        r32[k] = 0.0f;
        r64[k] = 0.0;
    }
    for (int j0 = 0; j0 < ncol; j0 += 64) {
        const int jn = min(j0 + 64, ncol);
        for (int jj = j0; jj < jn; jj++) {
            const float2 cj = cols[col0 + jj];
            float csum = 0.0f;
#pragma unroll
            for (int k = 0; k < kHuRows; k++) {
                const float q = fabsf(si[k] - cj.y) * __builtin_amdgcn_rcpf(si[k] + cj.y);
                const float pw = __builtin_amdgcn_exp2f(fmaf(-kLog2_100, q, kLog2_100));  // 100^(1 - q)
                const float a = fabsf(cj.x - hi[k]), b = pw - 1.0f;
                r32[k] = fmaf(a, b, r32[k]);
                if (!diag) csum = (FULL || row0 + tid + 256 * k < N) ? fmaf(a, b, csum) : csum;
            }
            if (!diag) {
#pragma unroll
                for (int m = 32; m >= 1; m >>= 1) csum += __shfl_xor(csum, m);
                if (lane == 0) colacc[wave][jj] = csum;
            }
        }
#pragma unroll
        for (int k = 0; k < kHuRows; k++) {
            r64[k] += (double)r32[k];
            r32[k] = 0.0f;
        }
    }
#pragma unroll
    for (int k = 0; k < kHuRows; k++) {
        const int r = row0 + tid + 256 * k;
        if (FULL || r < N) part[(size_t)c * N + r] = r64[k];
    }
    if (!diag) {
        __syncthreads();
        for (int jj = tid; jj < ncol; jj += 256)
            part[(size_t)p * N + col0 + jj] =
                (((double)colacc[0][jj] + (double)colacc[1][jj]) + (double)colacc[2][jj]) + (double)colacc[3][jj];
    }
}

__global__ __launch_bounds__(256) void k_hu_pairs_sym(const float2 *__restrict__ cols, int N,
                                                      double *__restrict__ part)
{
    __shared__ float colacc[4][kHuPanel];
    // (p, c) from the chunk-major triangular index q = c(c+1)/2 + p
    const int q = blockIdx.x;
    int c = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
    while ((c + 1) * (c + 2) / 2 <= q) c++;
    while (c * (c + 1) / 2 > q) c--;
    const int p = q - c * (c + 1) / 2;
    if ((p + 1) * kHuPanel <= N) hu_sym_panel<true>(cols, N, p, c, part, colacc);
    else hu_sym_panel<false>(cols, N, p, c, part, colacc);
}

static int hu_invariants_launch(sred_ctx *ctx, const Geometry &g, const float *d_img, const float *d_means,
                                const float *d_stds, float *d_hu_map, float2 *d_cols)
{
    float xavg = 0.0f, yavg = 0.0f;  // kernelGetHuMap.cl:147-157
    for (int i = 1; i <= g.bw; i++) xavg += (float)i;
    xavg /= (float)g.bw;
    for (int i = 1; i <= g.bh; i++) yavg += (float)i;
    yavg /= (float)g.bh;
    const double eps = java_float_literal(kEps);
    const int tw = kPmTile + 2 * g.brw, th = kPmTile + 2 * g.brh;
    dim3 grid((g.w + kPmTile - 1) / kPmTile, (g.h + kPmTile - 1) / kPmTile);
    hipLaunchKernelGGL(k_hu_invariants, grid, dim3(kPmTile * kPmTile), sizeof(float) * tw * th, ctx->stream, d_img,
                       d_means, d_stds, g.w, g.h, g.bw, g.bh, xavg, yavg, eps, d_hu_map, d_cols);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

static int interior_count(const Geometry &g, int *iw, int *ih)
{
    *iw = g.w - 2 * g.brw;
    *ih = g.h - 2 * g.brh;
    if (*iw <= 0 || *ih <= 0) {
        set_error("image smaller than the block");
        return SRED_ERR_ARG;
    }
    return SRED_OK;
}

int hu_invariants_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, const float *d_means, float *d_out)
{
    int iw, ih;
    SRED_TRY(interior_count(g, &iw, &ih));
    return hu_invariants_launch(ctx, g, d_img, d_means, nullptr, d_out, nullptr);
}

int global_hu_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, const float *d_means, const float *d_stds,
                  const int32_t *rows, int64_t nrows, float *d_out)
{
    int iw, ih;
    SRED_TRY(interior_count(g, &iw, &ih));
    const int ncols = iw * ih;
    PairRows pr;
    SRED_TRY(setup_rows(ctx, g, rows, nrows, ctx->pm_rows, &pr));
    SRED_TRY(zero_dev(ctx, d_out, sizeof(float) * g.n()));
    if (pr.R == 0) return SRED_OK;
    SRED_TRY(ctx->pm_cols.ensure(sizeof(float2) * (size_t)ncols));
    float2 *d_cols = ctx->pm_cols.as<float2>();
    SRED_TRY(hu_invariants_launch(ctx, g, d_img, d_means, d_stds, nullptr, d_cols));
    // the full map: each unordered panel pair once, while its partial sums
    // (npanels x N doubles) stay within the context's partial-sum budget
    const int64_t npan = (ncols + kHuPanel - 1) / kHuPanel;
    if (!pr.d_rowpix && npan * ncols * (int64_t)sizeof(double) <= (ctx->partial_mb << 20)) {
        SRED_TRY(ctx->pm_part.ensure(sizeof(double) * (size_t)npan * ncols));
        double *d_part = ctx->pm_part.as<double>();
        SRED_TRY(time_begin(ctx));
        hipLaunchKernelGGL(k_hu_pairs_sym, dim3((unsigned)(npan * (npan + 1) / 2)), dim3(256), 0, ctx->stream, d_cols,
                           ncols, d_part);
        SRED_HIP_CHECK(hipGetLastError());
        SRED_TRY(time_end(ctx, (int64_t)pr.R * ncols));
        hipLaunchKernelGGL(k_pm_finish, dim3((pr.R + 255) / 256), dim3(256), 0, ctx->stream, d_part, pr.R, (int)npan,
                           pr.d_rowpix, iw, g.w, g.brw, g.brh, 1.0 / 99.0, d_out);
        SRED_HIP_CHECK(hipGetLastError());
        return SRED_OK;
    }
    plan_chunks(&pr, ncols, 16384);
    SRED_TRY(ctx->pm_part.ensure(sizeof(double) * (size_t)pr.nchunks * pr.R));
    double *d_part = ctx->pm_part.as<double>();
    SRED_TRY(time_begin(ctx));
    hipLaunchKernelGGL(k_hu_pairs, dim3((pr.R + 255) / 256, pr.nchunks), dim3(256), 0, ctx->stream, d_cols, ncols,
                       pr.d_rowpix, pr.R, iw, g.w, g.brw, g.brh, pr.chunk, d_part);
    SRED_HIP_CHECK(hipGetLastError());
    SRED_TRY(time_end(ctx, (int64_t)pr.R * ncols));
    hipLaunchKernelGGL(k_pm_finish, dim3((pr.R + 255) / 256), dim3(256), 0, ctx->stream, d_part, pr.R, pr.nchunks,
                       pr.d_rowpix, iw, g.w, g.brw, g.brh, 1.0 / 99.0, d_out);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

// ---------------------------------------------------------------------------
// Phase correlation map
// ---------------------------------------------------------------------------

// Inverse-DFT twiddles (kernel argument): the row pass cos/sin(2 pi q / bw)
// and the column pass 2 cos / 2 sin(2 pi q / bh).  The forward table
// cos/sin(2 pi q / n), n = bw*bh, lives in device memory.
struct PhaseTw {
    double cw[kMaxPhaseSide], sw[kMaxPhaseSide], ch2[kMaxPhaseSide], sh2[kMaxPhaseSide];
};

// Unit spectrum of every interior block on the half-plane ky <= bh/2, one
// thread per (block, frequency): F(k) summed over the rectangle in raster
// order (angle index q = ((kx x mod bw) bh + (ky y mod bh) bw) mod n), then
// U = F / (|F| + eps).  spec[j * NT + m] (m = ky * bw + kx) = (Re U, Im U);
// sig[j] = std.  No contraction: U equals the CPU restatement's bit for bit.
template <int BW, int BH>
__global__ __launch_bounds__(256) void k_phase_spectra(const float *__restrict__ img, const float *__restrict__ means,
                                                       const float *__restrict__ stds, int w, int iw, int ncols,
                                                       const double *__restrict__ twc, const double *__restrict__ tws,
                                                       double eps, double2 *__restrict__ spec,
                                                       float *__restrict__ sig)
{
#pragma clang fp contract(off)
    constexpr int BRW = BW / 2, BRH = BH / 2, NT = BW * (BRH + 1), N = BW * BH;
    const int64_t t = (int64_t)blockIdx.x * 256 + threadIdx.x;
    if (t >= (int64_t)ncols * NT) return;
    const int j = (int)(t / NT), m = (int)(t - (int64_t)j * NT);
    const int ky = m / BW, kx = m - ky * BW;
    const int yj = j / iw, xj = j - yj * iw;  // top-left corner of the block = interior coords
    const int64_t p = (int64_t)(yj + BRH) * w + (xj + BRW);
    const double mean = (double)means[p], den = (double)stds[p] + eps;
    double a = 0.0, b = 0.0;
    for (int y = 0; y < BH; y++)
        for (int x = 0; x < BW; x++) {
            const int q = (((kx * x) % BW) * BH + ((ky * y) % BH) * BW) % N;
            const double v = ((double)img[(int64_t)(yj + y) * w + (xj + x)] - mean) / den;
            a += v * twc[q];
            b -= v * tws[q];
        }
    const double inv = 1.0 / (sqrt(a * a + b * b) + eps);
    spec[t] = make_double2(a * inv, b * inv);
    if (m == 0) sig[j] = stds[p];
}

// Row-major (SoA) copy of the rows' spectra, rspec[m * R + r], and their stds.
template <int NT>
__global__ __launch_bounds__(256) void k_phase_rows(const double2 *__restrict__ spec, const float *__restrict__ sig,
                                                    const int *__restrict__ rowpix, int R, int iw, int w, int brw,
                                                    int brh, double2 *__restrict__ rspec, float *__restrict__ rsig)
{
    const int r = blockIdx.x * 256 + threadIdx.x;
    if (r >= R) return;
    const int j = interior_index(row_pixel(rowpix, r, iw, w, brw, brh), iw, w, brw, brh);
    for (int m = 0; m < NT; m++) rspec[(size_t)m * R + r] = spec[(size_t)j * NT + m];
    rsig[r] = sig[j];
}

// The pair kernel.  Lane = row; the column's unit spectrum and std are
// wave-uniform.  UREG: the row spectrum is held in registers (small blocks)
// or read per pair from rspec (coalesced over lanes).  Products and sums are
// contracted to FMAs (r only decides the peak; the weight is exact either way).
// Displacement |d| of a pair from the half-plane spectra of its lower (lo)
// and higher (hi) interior index, given as functions of m returning (re, im):
// R = U_lo conj(U_hi), r = Re IDFT(R), the first strict positive maximum of
// r in raster order, wrapped to a signed displacement.
template <int BW, int BH, class LO, class HI>
__device__ __forceinline__ double phase_dist(LO &&lo_at, HI &&hi_at, const PhaseTw &tw)
{
    constexpr int BRW = BW / 2, BRH = BH / 2, NT = BW * (BRH + 1);
    double Rr[NT], Ri[NT];
#pragma unroll
    for (int m = 0; m < NT; m++) {
        const double2 u = lo_at(m);
        const double2 c = hi_at(m);
        Rr[m] = u.x * c.x + u.y * c.y;
        Ri[m] = u.y * c.x - u.x * c.y;
    }
    double best = 0.0;
    int px = 0, py = 0;
#pragma unroll
    for (int y = 0; y < BH; y++) {
        double Vr[BW], Vi[BW];
#pragma unroll
        for (int kx = 0; kx < BW; kx++) {
            double vr = Rr[kx], vi = Ri[kx];
#pragma unroll
            for (int ky = 1; ky <= BRH; ky++) {
                const int q = (ky * y) % BH;
                const double rr = Rr[ky * BW + kx], ri = Ri[ky * BW + kx];
                if (q == 0) {
                    vr += 2.0 * rr;
                    vi += 2.0 * ri;
                } else {
                    vr += rr * tw.ch2[q] - ri * tw.sh2[q];
                    vi += rr * tw.sh2[q] + ri * tw.ch2[q];
                }
            }
            Vr[kx] = vr;
            Vi[kx] = vi;
        }
#pragma unroll
        for (int x = 0; x < BW; x++) {
            double v = Vr[0];
#pragma unroll
            for (int kx = 1; kx < BW; kx++) {
                const int q = (kx * x) % BW;
                if (q == 0)
                    v += Vr[kx];
                else
                    v += Vr[kx] * tw.cw[q] - Vi[kx] * tw.sw[q];
            }
            if (v > best) {
                best = v;
                px = x;
                py = y;
            }
        }
    }
    const int dx = px <= BRW ? px : px - BW, dy = py <= BRH ? py : py - BH;
    return sqrt((double)(dx * dx + dy * dy));
}

// kernelGetPhaseCorrelationMap.cl's weight (:167-178), symmetric in (s_i, s_j)
__device__ __forceinline__ float phase_weight(float si, float sj)
{
    return (si == sj) ? 1.0f : 1.0f - (2.0f * (fabsf(si - sj) / (fabsf(si) + fabsf(sj))));
}

template <int BW, int BH, bool UREG>
__global__ __launch_bounds__(64) void k_phase_pairs(const double2 *__restrict__ spec, const float *__restrict__ sig,
                                                    int ncols, const double2 *__restrict__ rspec,
                                                    const float *__restrict__ rsig, int R, int chunk, PhaseTw tw,
                                                    const int *__restrict__ rowpix, int iw, int w, int brw, int brh,
                                                    double *__restrict__ part)
{
    constexpr int BRH = BH / 2, NT = BW * (BRH + 1);
    const int r = blockIdx.x * 64 + threadIdx.x;
    if (r >= R) return;
    const int ri = interior_index(row_pixel(rowpix, r, iw, w, brw, brh), iw, w, brw, brh);
    double ur[UREG ? NT : 1], ui[UREG ? NT : 1];
    if constexpr (UREG) {
#pragma unroll
        for (int m = 0; m < NT; m++) {
            const double2 u = rspec[(size_t)m * R + r];
            ur[m] = u.x;
            ui[m] = u.y;
        }
    }
    const float si = rsig[r];
    const int c0 = blockIdx.y * chunk, c1 = min(c0 + chunk, ncols);
    auto row_at = [&](int m) -> double2 {
        if constexpr (UREG) return make_double2(ur[m], ui[m]);
        else return rspec[(size_t)m * R + r];
    };
    double acc = 0.0;
    for (int j = c0; j < c1; j++) {
        const double2 *uj = spec + (size_t)j * NT;
        // the pair's lower interior index leads (per lane: j < ri swaps the roles)
        const bool swap = j < ri;
        const double dist = phase_dist<BW, BH>([&](int m) { return swap ? uj[m] : row_at(m); },
                                               [&](int m) { return swap ? row_at(m) : uj[m]; }, tw);
        acc += dist * (double)phase_weight(si, sig[j]);
    }
    part[(size_t)blockIdx.y * R + r] = acc;
}

// Symmetric form of k_phase_pairs for the full map: the pair's displacement
// is taken with its lower interior index first, so each unordered pair has
// one term, evaluated once here.  One workgroup per pair of 256-pixel panels
// (p, c), c >= p, thread t holding row p P + t; per column j of panel c
// (wave-uniform) the thread adds the term to its row and the wave reduces
// it (butterfly, FP64) into its own LDS row of column sums.  Off the
// diagonal every row index is below every column index; on it only pairs
// i < j are taken.  Row totals go to part[c][i] and the four waves' column
// totals to part[p][j] (on the diagonal both to part[p][i]): one writer per
// slot, summed in slot order by k_pm_finish.
constexpr int kPhPanel = 256;
template <int BW, int BH, bool UREG>
__global__ __launch_bounds__(256) void k_phase_pairs_sym(const double2 *__restrict__ spec,
                                                         const float *__restrict__ sig, int N,
                                                         const double2 *__restrict__ rspec, PhaseTw tw,
                                                         double *__restrict__ part)
{
    constexpr int NT = BW * (BH / 2 + 1);
    __shared__ double colacc[4][kPhPanel];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    // (p, c) from the chunk-major triangular index q = c(c+1)/2 + p
    const int q = blockIdx.x;
    int c = (int)((sqrtf(8.0f * (float)q + 1.0f) - 1.0f) * 0.5f);
    while ((c + 1) * (c + 2) / 2 <= q) c++;
    while (c * (c + 1) / 2 > q) c--;
    const int p = q - c * (c + 1) / 2;
    const bool diag = p == c;
    const int i = p * kPhPanel + tid;
    const bool live = i < N;
    double ur[UREG ? NT : 1], ui[UREG ? NT : 1];
    if constexpr (UREG) {
#pragma unroll
        for (int m = 0; m < NT; m++) {
            const double2 u = live ? rspec[(size_t)m * N + i] : make_double2(0.0, 0.0);
            ur[m] = u.x;
            ui[m] = u.y;
        }
    }
    auto row_at = [&](int m) -> double2 {
        if constexpr (UREG) return make_double2(ur[m], ui[m]);
        else return rspec[(size_t)m * N + i];
    };
    const float si = live ? sig[i] : 0.0f;
    const int col0 = c * kPhPanel, ncol = min(kPhPanel, N - col0);
    double racc = 0.0;
    for (int jj = 0; jj < ncol; jj++) {
        const int j = col0 + jj;
        const double2 *uj = spec + (size_t)j * NT;
        double t = 0.0;
        if (live && (!diag || i < j))
            t = phase_dist<BW, BH>(row_at, [&](int m) { return uj[m]; }, tw) * (double)phase_weight(si, sig[j]);
        racc += t;
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) t += __shfl_xor(t, m);
        if (lane == 0) colacc[wave][jj] = t;
    }
    __syncthreads();
    if (diag) {
        if (live)
            part[(size_t)p * N + i] =
                racc + (((colacc[0][tid] + colacc[1][tid]) + colacc[2][tid]) + colacc[3][tid]);
    } else {
        if (live) part[(size_t)c * N + i] = racc;
        for (int jj = tid; jj < ncol; jj += 256)
            part[(size_t)p * N + col0 + jj] = ((colacc[0][jj] + colacc[1][jj]) + colacc[2][jj]) + colacc[3][jj];
    }
}

struct PhaseLaunch {
    const float *img, *means, *stds;
    int w, iw, ncols;
    double eps;
    const double *twc, *tws;
    double2 *spec, *rspec;
    float *sig, *rsig;
    PairRows pr;
    PhaseTw tw;
    double *part;
    int brw, brh;
    int npan;  // > 0: the symmetric kernel over npan panels (part is [npan][ncols])
};

template <int BW, int BH>
static int phase_run(sred_ctx *ctx, const PhaseLaunch &L, int brw, int brh)
{
    constexpr int NT = BW * (BH / 2 + 1);
    constexpr bool UREG = NT <= 15;
    const int64_t nt = (int64_t)L.ncols * NT;
    hipLaunchKernelGGL((k_phase_spectra<BW, BH>), dim3((unsigned)((nt + 255) / 256)), dim3(256), 0, ctx->stream,
                       L.img, L.means, L.stds, L.w, L.iw, L.ncols, L.twc, L.tws, L.eps, L.spec, L.sig);
    hipLaunchKernelGGL((k_phase_rows<NT>), dim3((L.pr.R + 255) / 256), dim3(256), 0, ctx->stream, L.spec, L.sig,
                       L.pr.d_rowpix, L.pr.R, L.iw, L.w, brw, brh, L.rspec, L.rsig);
    SRED_HIP_CHECK(hipGetLastError());
    SRED_TRY(time_begin(ctx));
    if (L.npan > 0)
        hipLaunchKernelGGL((k_phase_pairs_sym<BW, BH, UREG>), dim3((unsigned)((int64_t)L.npan * (L.npan + 1) / 2)),
                           dim3(256), 0, ctx->stream, L.spec, L.sig, L.ncols, L.rspec, L.tw, L.part);
    else
        hipLaunchKernelGGL((k_phase_pairs<BW, BH, UREG>), dim3((L.pr.R + 63) / 64, L.pr.nchunks), dim3(64), 0,
                           ctx->stream, L.spec, L.sig, L.ncols, L.rspec, L.rsig, L.pr.R, L.pr.chunk, L.tw,
                           L.pr.d_rowpix, L.iw, L.w, L.brw, L.brh, L.part);
    SRED_HIP_CHECK(hipGetLastError());
    SRED_TRY(time_end(ctx, (int64_t)L.pr.R * L.ncols));
    return SRED_OK;
}

template <int BW>
static int phase_dispatch_h(sred_ctx *ctx, const PhaseLaunch &L, int bh, int brw, int brh)
{
    switch (bh) {
        case 3: return phase_run<BW, 3>(ctx, L, brw, brh);
        case 5: return phase_run<BW, 5>(ctx, L, brw, brh);
        case 7: return phase_run<BW, 7>(ctx, L, brw, brh);
        case 9: return phase_run<BW, 9>(ctx, L, brw, brh);
        case 11: return phase_run<BW, 11>(ctx, L, brw, brh);
    }
    set_error("phase correlation supports odd block sides 3..11");
    return SRED_ERR_ARG;
}

int phase_map_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, const float *d_means, const float *d_stds,
                  const int32_t *rows, int64_t nrows, float *d_out)
{
    if (g.bw > kMaxPhaseSide || g.bh > kMaxPhaseSide) {
        set_error("phase correlation supports odd block sides 3..11");
        return SRED_ERR_ARG;
    }
    int iw, ih;
    SRED_TRY(interior_count(g, &iw, &ih));
    PhaseLaunch L{};
    L.img = d_img;
    L.means = d_means;
    L.stds = d_stds;
    L.w = g.w;
    L.iw = iw;
    L.ncols = iw * ih;
    L.eps = java_float_literal(kEps);
    SRED_TRY(setup_rows(ctx, g, rows, nrows, ctx->pm_rows, &L.pr));
    SRED_TRY(zero_dev(ctx, d_out, sizeof(float) * g.n()));
    if (L.pr.R == 0) return SRED_OK;
    const int n = g.bw * g.bh, NT = g.bw * (g.brh + 1);
    // twiddles (host libm, as the CPU restatement)
    const double two_pi = 6.283185307179586476925;
    std::vector<double> tab(2 * (size_t)n);
    for (int q = 0; q < n; q++) {
        tab[q] = std::cos(two_pi * (double)q / (double)n);
        tab[n + q] = std::sin(two_pi * (double)q / (double)n);
    }
    for (int q = 0; q < g.bw; q++) {
        L.tw.cw[q] = std::cos(two_pi * (double)q / (double)g.bw);
        L.tw.sw[q] = std::sin(two_pi * (double)q / (double)g.bw);
    }
    for (int q = 0; q < g.bh; q++) {
        L.tw.ch2[q] = 2.0 * std::cos(two_pi * (double)q / (double)g.bh);
        L.tw.sh2[q] = 2.0 * std::sin(two_pi * (double)q / (double)g.bh);
    }
    plan_chunks(&L.pr, L.ncols, 4096);
    L.brw = g.brw;
    L.brh = g.brh;
    // the full map: each unordered panel pair once, while its partial sums
    // (npanels x ncols doubles) stay within the context's partial-sum budget
    const int64_t npan = (L.ncols + kPhPanel - 1) / kPhPanel;
    L.npan = (!L.pr.d_rowpix && npan * L.ncols * (int64_t)sizeof(double) <= (ctx->partial_mb << 20)) ? (int)npan : 0;
    const int nslots = L.npan > 0 ? L.npan : L.pr.nchunks;
    // scratch: [twiddles | spec | rspec | part | sig | rsig]
    const size_t b_tw = sizeof(double) * 2 * n;
    const size_t b_spec = sizeof(double2) * (size_t)L.ncols * NT;
    const size_t b_rspec = sizeof(double2) * (size_t)L.pr.R * NT;
    const size_t b_part = sizeof(double) * (size_t)nslots * L.pr.R;
    const size_t b_sig = sizeof(float) * ((size_t)L.ncols + 64), b_rsig = sizeof(float) * ((size_t)L.pr.R + 64);
    auto al = [](size_t b) { return (b + 255) / 256 * 256; };
    SRED_TRY(ctx->pm_spec.ensure(al(b_tw) + al(b_spec) + al(b_rspec) + al(b_part) + al(b_sig) + al(b_rsig)));
    char *base = ctx->pm_spec.as<char>();
    double *d_tw = (double *)base;
    base += al(b_tw);
    L.spec = (double2 *)base;
    base += al(b_spec);
    L.rspec = (double2 *)base;
    base += al(b_rspec);
    L.part = (double *)base;
    base += al(b_part);
    L.sig = (float *)base;
    base += al(b_sig);
    L.rsig = (float *)base;
    SRED_HIP_CHECK(hipMemcpyAsync(d_tw, tab.data(), b_tw, hipMemcpyHostToDevice, ctx->stream));
    L.twc = d_tw;
    L.tws = d_tw + n;
    int rc = SRED_ERR_ARG;
    switch (g.bw) {
        case 3: rc = phase_dispatch_h<3>(ctx, L, g.bh, g.brw, g.brh); break;
        case 5: rc = phase_dispatch_h<5>(ctx, L, g.bh, g.brw, g.brh); break;
        case 7: rc = phase_dispatch_h<7>(ctx, L, g.bh, g.brw, g.brh); break;
        case 9: rc = phase_dispatch_h<9>(ctx, L, g.bh, g.brw, g.brh); break;
        case 11: rc = phase_dispatch_h<11>(ctx, L, g.bh, g.brw, g.brh); break;
        default: set_error("phase correlation supports odd block sides 3..11");
    }
    if (rc != SRED_OK) {
        (void)hipStreamSynchronize(ctx->stream);  // tab must outlive the copy
        return rc;
    }
    hipLaunchKernelGGL(k_pm_finish, dim3((L.pr.R + 255) / 256), dim3(256), 0, ctx->stream, L.part, L.pr.R, nslots,
                       L.pr.d_rowpix, iw, g.w, g.brw, g.brh, 1.0, d_out);
    SRED_HIP_CHECK(hipGetLastError());
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));  // tab must outlive the copy
    return SRED_OK;
}

}  // namespace sred
