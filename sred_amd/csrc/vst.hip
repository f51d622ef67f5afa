// vst.hip -- variance stabilisation before the path: the generalised
// Anscombe transform (GAT) and the simplex fit of its parameters.
//
// Reference: Utils.getInputImage2D (Utils.java:693-703) runs
// GATMinimizer2D(image, w, h, 1, 10, 100, maxIter) (GATMinimizer2D.java:33-61,
// ImageJ's Minimizer over userFunction :64-110), casts the fitted gain, sigma
// and offset to float and applies VarianceStabilisingTransform2D_.getGAT
// (VarianceStabilisingTransform2D_.java:136-151).
//
//   * getGAT: v <- (g v + ref < 0) ? 0 : (2/g) sqrt(g v + ref),
//     ref = 3/8 g^2 + s^2 - g o, in double, rounded to float.  k_gat is an
//     elementwise HBM-bound kernel (4 B read + 4 B written per pixel).
//   * userFunction: the image in nbx x nby blocks of (w/6) x (h/6) pixels;
//     per block the mean and variance of the transformed pixels (the fit's
//     own transform, special case v <= -ref/g); error = sum (var - 1)^2 / nb.
//     k_gat_rows takes one (image row, block column) per wave for all
//     parameter sets of the launch (one pixel load, up to four transforms):
//     lane l adds the transformed pixels l, l + 64, ... of the block row (and
//     their float squares) in double, then a fixed xor butterfly over the 64
//     lanes gives the row sums.  k_gat_error (one workgroup per parameter
//     set) combines the row sums of each block the same way (row r to lane
//     r mod 64) and then the block terms (block b to lane b mod 64), and
//     writes the value to pinned host memory.  An evaluation is deterministic
//     and equals the CPU restatement bit for bit (IEEE double arithmetic,
//     commutative adds and correctly rounded square roots on both).
//   * The minimiser: ImageJ's Minimizer (restarts, random initial simplex)
//     is third-party and not restated; this is a Nelder-Mead simplex with
//     the Lagarias et al. (1998) coefficients, initial simplex x0 + d_i e_i
//     with d = (1, 10, 100), stopping after max_iter iterations or when all
//     vertices have the same objective (the reference's setMaxError(0)); a
//     NaN objective (invalid parameters, e.g. gain 0) ranks worst, as
//     ImageJ's Minimizer treats NaN.
//     Each iteration evaluates the reflection, expansion and both
//     contractions in one launch (four parameter sets) and then takes the
//     branch the values select, so the path is the sequential algorithm's.
//   * Volumes (GATMinimizer3D.java:79-157, run by the
//     VarianceStabilisingTransform3D_ plugin): the same kernels on rows
//     addressed by (slice, row), blocks of bw x bh x bd voxels whose rows are
//     taken slice by slice, squares in double (Utils.getMeanAndVarBlock3D),
//     NaN for gain <= 0, sigma < 0 or offset < 0, and the 3D block grid
//     (sides max(n/6, 1), halved along axes with fewer than 2 blocks).
#include <algorithm>
#include <cmath>
#include <vector>

#include "sred_internal.h"

namespace sred {

constexpr int kGatSets = 4;        // parameter sets per objective launch
constexpr int kGatErrThreads = 1024;  // k_gat_error workgroup (16 waves)
constexpr int kGatMaxBlocks = 4096;   // blocks of the objective (LDS table of block terms)

struct GatSets {
    double gain[kGatSets], ref[kGatSets];
};

__global__ __launch_bounds__(256) void k_gat(float *__restrict__ px, int64_t n, double gain, double ref)
{
#pragma clang fp contract(off)
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = (double)px[i];
        if ((gain * v + ref) < 0.0)
            v = 0.0;
        else
            v = (2.0 / gain) * sqrt(gain * v + ref);
        px[i] = (float)v;
    }
}

// Sum over the 64 lanes of a wave: for s = 32, ..., 1 each lane adds lane
// (l ^ s)'s value; every lane ends with the same result.
__device__ __forceinline__ double lane_tree64(double a)
{
#pragma clang fp contract(off)
#pragma unroll
    for (int s = 32; s >= 1; s >>= 1) a = a + __shfl_xor(a, s, 64);
    return a;
}

// rows[set][j][bx] = (sum v, sum v*v) over row j of block column bx; one
// wave per (j, bx) evaluates every parameter set of the launch.  Row j is
// row j mod rps of slice j / rps (plane voxels apart); DSQ squares in double
// (the volume objective), else in float (the image objective).
//
// Rows of at most 32 elements share a wave: segments of S = 2^seg_log2
// lanes (the smallest power of 2 >= bw) each take one row, and the butterfly
// runs within the segment.  That equals lane_tree64 over the row alone: the
// steps across segments would add exact zeros (the transformed values and
// their squares are >= 0).
template <bool DSQ>
__global__ __launch_bounds__(256) void k_gat_rows(const float *__restrict__ px, int w, int bw, int nbx, int nrows,
                                                  int rps, int64_t plane, int nsets, GatSets ps, int seg_log2,
                                                  double2 *__restrict__ rows)
{
#pragma clang fp contract(off)
    const int lane = threadIdx.x & 63;
    const int S = 1 << seg_log2, li = lane & (S - 1);
    const int wave_id = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6);
    const int t = (wave_id << (6 - seg_log2)) + (lane >> seg_log2);
    const int per_set = nrows * nbx;
    if ((wave_id << (6 - seg_log2)) >= per_set) return;  // wave-uniform: no row of this wave exists
    const bool valid = t < per_set;
    const int tt = valid ? t : 0;
    const int j = tt / nbx, bx = tt - j * nbx;
    const int z = j / rps, y = j - z * rps;
    double lim[kGatSets], two_g[kGatSets], rm[kGatSets], rs[kGatSets];
#pragma unroll
    for (int s = 0; s < kGatSets; s++) {
        lim[s] = -ps.ref[s] / ps.gain[s];
        two_g[s] = 2.0 / ps.gain[s];
        rm[s] = rs[s] = 0.0;
    }
    const float *p = px + z * plane + (int64_t)y * w + (int64_t)bx * bw;
    for (int i = li; i < (valid ? bw : 0); i += S) {
        const double d = (double)p[i];
#pragma unroll
        for (int s = 0; s < kGatSets; s++) {
            if (s < nsets) {
                const float v = d <= lim[s] ? 0.0f : (float)(two_g[s] * sqrt(ps.gain[s] * d + ps.ref[s]));
                rm[s] += v;
                if (DSQ) rs[s] += (double)v * (double)v;
                else rs[s] += v * v;
            }
        }
    }
#pragma unroll
    for (int s = 0; s < kGatSets; s++) {
        if (s < nsets) {
            double a = rm[s], b = rs[s];
            for (int m = S >> 1; m >= 1; m >>= 1) {
                a = a + __shfl_xor(a, m, 64);
                b = b + __shfl_xor(b, m, 64);
            }
            if (li == 0 && valid) rows[(size_t)s * per_set + t] = make_double2(a, b);
        }
    }
}

// err[set] from the row sums of the set (one workgroup per set): each wave
// takes blocks wave, wave + 16, ... (z, y, x order; a block's bh * bd rows
// slice by slice, row k to lane k mod 64); the block terms (block b to lane
// b mod 64) are combined by wave 0.
__global__ __launch_bounds__(kGatErrThreads) void k_gat_error(const double2 *__restrict__ rows, int bw, int bh,
                                                              int bd, int nbx, int nby, int nbz, int nrows, int rps,
                                                              double *__restrict__ err)
{
#pragma clang fp contract(off)
    __shared__ double be[kGatMaxBlocks];
    const int set = blockIdx.x, lane = threadIdx.x & 63, wave = threadIdx.x >> 6, nb = nbx * nby * nbz;
    const double2 *r = rows + (size_t)set * nrows * nbx;
    for (int b = wave; b < nb; b += kGatErrThreads / 64) {
        const int bz = b / (nbx * nby), rem = b - bz * (nbx * nby), by = rem / nbx, bx = rem - by * nbx;
        double mean = 0.0, sq = 0.0;
        for (int k = lane; k < bh * bd; k += 64) {
            const int kz = k / bh, ky = k - kz * bh;
            const double2 q = r[((size_t)(bz * bd + kz) * rps + by * bh + ky) * nbx + bx];
            mean += q.x;
            sq += q.y;
        }
        mean = lane_tree64(mean);
        sq = lane_tree64(sq);
        const double n = (double)((int64_t)bw * bh * bd);
        mean = mean / n;
        const double var = sq / n - mean * mean;
        const double delta = var - 1.0;
        if (lane == 0) be[b] = (delta * delta) / (double)nb;
    }
    __syncthreads();
    if (wave == 0) {
        double e = 0.0;
        for (int q = lane; q < nb; q += 64) e += be[q];
        e = lane_tree64(e);
        if (lane == 0) err[set] = e;
    }
}

// Volumes: the same block terms, one block per wave over grid.y workgroups
// per set, into bterm[set][b]; k_gat_error_final then combines them in
// k_gat_error's order (block b to lane b mod 64, then lane_tree64).
__global__ __launch_bounds__(kGatErrThreads) void k_gat_error_part(const double2 *__restrict__ rows, int bw, int bh,
                                                                   int bd, int nbx, int nby, int nbz, int nrows,
                                                                   int rps, double *__restrict__ bterm)
{
#pragma clang fp contract(off)
    const int set = blockIdx.x, lane = threadIdx.x & 63, nb = nbx * nby * nbz;
    const int b = blockIdx.y * (kGatErrThreads / 64) + (threadIdx.x >> 6);
    if (b >= nb) return;
    const double2 *r = rows + (size_t)set * nrows * nbx;
    const int bz = b / (nbx * nby), rem = b - bz * (nbx * nby), by = rem / nbx, bx = rem - by * nbx;
    double mean = 0.0, sq = 0.0;
    for (int k = lane; k < bh * bd; k += 64) {
        const int kz = k / bh, ky = k - kz * bh;
        const double2 q = r[((size_t)(bz * bd + kz) * rps + by * bh + ky) * nbx + bx];
        mean += q.x;
        sq += q.y;
    }
    mean = lane_tree64(mean);
    sq = lane_tree64(sq);
    const double n = (double)((int64_t)bw * bh * bd);
    mean = mean / n;
    const double var = sq / n - mean * mean;
    const double delta = var - 1.0;
    if (lane == 0) bterm[(size_t)set * nb + b] = (delta * delta) / (double)nb;
}

__global__ __launch_bounds__(64) void k_gat_error_final(const double *__restrict__ bterm, int nb,
                                                        double *__restrict__ err)
{
#pragma clang fp contract(off)
    const int set = blockIdx.x, lane = threadIdx.x;
    double e = 0.0;
    for (int q = lane; q < nb; q += 64) e += bterm[(size_t)set * nb + q];
    e = lane_tree64(e);
    if (lane == 0) err[set] = e;
}

static double gat_ref(double gain, double sigma, double offset)
{
#pragma clang fp contract(off)
    return (3.0 / 8.0) * gain * gain + sigma * sigma - gain * offset;
}

int gat_ref_dev(sred_ctx *ctx, float *d_img, int64_t n, double gain, double ref)
{
    if (n <= 0) return SRED_OK;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_gat, dim3(blocks), dim3(256), 0, ctx->stream, d_img, n, gain, ref);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

int gat_dev(sred_ctx *ctx, float *d_img, int64_t n, double gain, double sigma, double offset)
{
    return gat_ref_dev(ctx, d_img, n, gain, gat_ref(gain, sigma, offset));
}

namespace {

// One device evaluation of up to kGatSets parameter sets (synchronises).
struct GatEval {
    sred_ctx *ctx;
    const float *d_img;
    int w, h, bw, bh, nbx, nby, nrows;
    int bd = 1, nbz = 1, rps = 0;  // volumes: block depth, block slices, rows per slice
    int64_t plane = 0;             // voxels per slice (0 for an image)
    bool volume = false;           // double squares, NaN for invalid parameters
    double2 *d_rows;
    double *d_bterm = nullptr;     // volumes: per-block terms of every set
    double *h_err;  // pinned host memory written by k_gat_error
    int64_t launches = 0, pixel_evals = 0;
    double rows_ms = 0.0;  // k_gat_rows time summed over the launches (HIP events, when fit_timing is on)
    bool rank_nan_worst = false;  // the simplex's evaluations map NaN to +infinity

    int run(const double (*x)[3], int nsets, double *f)
    {
        GatSets ps{};
        for (int s = 0; s < nsets; s++) {
            ps.gain[s] = x[s][0];
            ps.ref[s] = gat_ref(x[s][0], x[s][1], x[s][2]);
        }
        int seg_log2 = 0;
        while ((1 << seg_log2) < bw && seg_log2 < 6) seg_log2++;
        const int waves = (nrows * nbx + (64 >> seg_log2) - 1) >> (6 - seg_log2);
        if (ctx->fit_timing) SRED_HIP_CHECK(hipEventRecord(ctx->ev0, ctx->stream));
        if (volume)
            hipLaunchKernelGGL(k_gat_rows<true>, dim3((waves + 3) / 4), dim3(256), 0, ctx->stream, d_img, w, bw, nbx,
                               nrows, rps, plane, nsets, ps, seg_log2, d_rows);
        else
            hipLaunchKernelGGL(k_gat_rows<false>, dim3((waves + 3) / 4), dim3(256), 0, ctx->stream, d_img, w, bw, nbx,
                               nrows, rps, plane, nsets, ps, seg_log2, d_rows);
        if (ctx->fit_timing) SRED_HIP_CHECK(hipEventRecord(ctx->ev1, ctx->stream));
        const int nb = nbx * nby * nbz;
        if (volume) {  // many blocks: spread over workgroups, then the fixed-order sum
            const int gy = (nb + kGatErrThreads / 64 - 1) / (kGatErrThreads / 64);
            hipLaunchKernelGGL(k_gat_error_part, dim3(nsets, gy), dim3(kGatErrThreads), 0, ctx->stream, d_rows, bw,
                               bh, bd, nbx, nby, nbz, nrows, rps, d_bterm);
            hipLaunchKernelGGL(k_gat_error_final, dim3(nsets), dim3(64), 0, ctx->stream, d_bterm, nb, h_err);
        } else {
            hipLaunchKernelGGL(k_gat_error, dim3(nsets), dim3(kGatErrThreads), 0, ctx->stream, d_rows, bw, bh, bd,
                               nbx, nby, nbz, nrows, rps, h_err);
        }
        SRED_HIP_CHECK(hipGetLastError());
        SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        for (int s = 0; s < nsets; s++) f[s] = h_err[s];
        if (volume)  // GATMinimizer3D.java:84-86
            for (int s = 0; s < nsets; s++)
                if (x[s][0] <= 0.0 || x[s][1] < 0.0 || x[s][2] < 0.0) f[s] = NAN;
        if (rank_nan_worst)  // invalid parameters (NaN objective) rank worst, as ImageJ's Minimizer treats NaN
            for (int s = 0; s < nsets; s++)
                if (f[s] != f[s]) f[s] = INFINITY;
        if (ctx->fit_timing) {
            float ms = 0.0f;
            SRED_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
            rows_ms += ms;
        }
        launches++;
        pixel_evals += (int64_t)nsets * nrows * nbx * bw;
        return SRED_OK;
    }
};

int gat_setup(sred_ctx *ctx, const float *d_img, int w, int h, GatEval *e)
{
    const int bw = w / 6, bh = h / 6;  // GATMinimizer2D.java:77-78
    if (bw <= 0 || bh <= 0) {
        set_error("the GAT fit needs an image of at least 6 x 6 pixels (the reference divides by zero below)");
        return SRED_ERR_ARG;
    }
    e->ctx = ctx;
    e->d_img = d_img;
    e->w = w;
    e->h = h;
    e->bw = bw;
    e->bh = bh;
    e->nbx = w / bw;
    e->nby = h / bh;
    e->nrows = e->nby * bh;
    e->rps = e->nrows;
    if (e->nbx * e->nby > kGatMaxBlocks) {
        set_error("unsupported GAT block grid");
        return SRED_ERR_ARG;
    }
    SRED_TRY(ctx->pm_part.ensure(sizeof(double2) * (size_t)kGatSets * e->nrows * e->nbx + 256));
    e->d_rows = ctx->pm_part.as<double2>();
    if (!ctx->h_fit) SRED_HIP_CHECK(hipHostMalloc((void **)&ctx->h_fit, sizeof(double) * kGatSets, hipHostMallocDefault));
    e->h_err = ctx->h_fit;
    SRED_TRY(time_begin(ctx));  // creates the events; the fit sums its launches itself
    return SRED_OK;
}

// The block grid of GATMinimizer3D.userFunction (GATMinimizer3D.java:108-127).
int gat_setup3d(sred_ctx *ctx, const float *d_vol, int w, int h, int d, GatEval *e)
{
    int a = std::max(w / 6, 1), b = std::max(h / 6, 1), c = std::max(d / 6, 1);
    int nx = w / a, ny = h / b, nz = d / c;
    while (nx < 2 || ny < 2 || nz < 2) {
        const int a0 = a, b0 = b, c0 = c;
        if (nx < 2 && a > 1) a = std::max(a / 2, 1);
        if (ny < 2 && b > 1) b = std::max(b / 2, 1);
        if (nz < 2 && c > 1) c = std::max(c / 2, 1);
        nx = w / a, ny = h / b, nz = d / c;
        if (a == 1 && b == 1 && c == 1) break;
        if (a == a0 && b == b0 && c == c0) {
            set_error("GATMinimizer3D's block search does not terminate for this volume (a 1-voxel axis while "
                      "another block side is above 1)");
            return SRED_ERR_ARG;
        }
    }
    if ((int64_t)nx * ny * nz > kGatMaxBlocks) {
        set_error("unsupported GAT block grid (more than 4096 blocks)");
        return SRED_ERR_ARG;
    }
    e->ctx = ctx;
    e->d_img = d_vol;
    e->w = w;
    e->h = h;
    e->bw = a;
    e->bh = b;
    e->bd = c;
    e->nbx = nx;
    e->nby = ny;
    e->nbz = nz;
    e->rps = ny * b;
    e->nrows = nz * c * e->rps;
    e->plane = (int64_t)w * h;
    e->volume = true;
    const size_t rows_bytes = (sizeof(double2) * (size_t)kGatSets * e->nrows * e->nbx + 255) / 256 * 256;
    SRED_TRY(ctx->pm_part.ensure(rows_bytes + sizeof(double) * (size_t)kGatSets * nx * ny * nz + 256));
    e->d_rows = ctx->pm_part.as<double2>();
    e->d_bterm = reinterpret_cast<double *>(ctx->pm_part.as<char>() + rows_bytes);
    if (!ctx->h_fit) SRED_HIP_CHECK(hipHostMalloc((void **)&ctx->h_fit, sizeof(double) * kGatSets, hipHostMallocDefault));
    e->h_err = ctx->h_fit;
    SRED_TRY(time_begin(ctx));
    return SRED_OK;
}

// Nelder-Mead as the CPU restatement (same expressions, no contraction), with
// the candidate points of an iteration evaluated together.
int gat_minimize_run(sred_ctx *ctx, GatEval &e, double gain0, double sigma0, double offset0, int max_iter,
                     double *params, double *error);

}  // namespace

int gat_objective3d_dev(sred_ctx *ctx, const float *d_vol, int w, int h, int d, double gain, double sigma,
                        double offset, double *error)
{
    GatEval e;
    SRED_TRY(gat_setup3d(ctx, d_vol, w, h, d, &e));
    const double x[1][3] = {{gain, sigma, offset}};
    SRED_TRY(e.run(x, 1, error));
    ctx->last_launches = 1;
    ctx->last_pairs = e.pixel_evals;
    ctx->summed_ms = (float)e.rows_ms;
    return SRED_OK;
}

int gat_minimize3d_dev(sred_ctx *ctx, const float *d_vol, int w, int h, int d, double gain0, double sigma0,
                       double offset0, int max_iter, double *params, double *error)
{
    GatEval e;
    SRED_TRY(gat_setup3d(ctx, d_vol, w, h, d, &e));
    return gat_minimize_run(ctx, e, gain0, sigma0, offset0, max_iter, params, error);
}

int gat_objective_dev(sred_ctx *ctx, const float *d_img, int w, int h, double gain, double sigma, double offset,
                      double *error)
{
    GatEval e;
    SRED_TRY(gat_setup(ctx, d_img, w, h, &e));
    const double x[1][3] = {{gain, sigma, offset}};
    SRED_TRY(e.run(x, 1, error));
    ctx->last_launches = 1;
    ctx->last_pairs = e.pixel_evals;
    ctx->summed_ms = (float)e.rows_ms;
    return SRED_OK;
}

int gat_minimize_dev(sred_ctx *ctx, const float *d_img, int w, int h, double gain0, double sigma0, double offset0,
                     int max_iter, double *params, double *error)
{
    GatEval e;
    SRED_TRY(gat_setup(ctx, d_img, w, h, &e));
    return gat_minimize_run(ctx, e, gain0, sigma0, offset0, max_iter, params, error);
}

namespace {

int gat_minimize_run(sred_ctx *ctx, GatEval &e, double gain0, double sigma0, double offset0, int max_iter,
                     double *params, double *error)
{
#pragma clang fp contract(off)
    e.rank_nan_worst = true;
    const double d[3] = {1.0, 10.0, 100.0};  // GATMinimizer2D.java:43-45
    double v[4][3], fv[4];
    for (int i = 0; i < 4; i++) {
        v[i][0] = gain0;
        v[i][1] = sigma0;
        v[i][2] = offset0;
        if (i > 0) v[i][i - 1] += d[i - 1];
    }
    SRED_TRY(e.run(v, 4, fv));
    for (int it = 0; it < max_iter; it++) {
        for (int i = 1; i < 4; i++)  // stable insertion sort by f
            for (int j = i; j > 0 && fv[j] < fv[j - 1]; j--) {
                std::swap(fv[j], fv[j - 1]);
                for (int k = 0; k < 3; k++) std::swap(v[j][k], v[j - 1][k]);
            }
        if (fv[3] == fv[0]) break;
        double c[3], pts[4][3], fp[4];  // reflection, expansion, outside and inside contraction
        for (int k = 0; k < 3; k++) c[k] = (v[0][k] + v[1][k] + v[2][k]) / 3.0;
        for (int k = 0; k < 3; k++) {
            pts[0][k] = c[k] + (c[k] - v[3][k]);
            pts[1][k] = c[k] + 2.0 * (c[k] - v[3][k]);
            pts[2][k] = c[k] + 0.5 * (pts[0][k] - c[k]);
            pts[3][k] = c[k] + 0.5 * (v[3][k] - c[k]);
        }
        SRED_TRY(e.run(pts, 4, fp));
        const double fr = fp[0];
        int take = -1;
        if (fr < fv[0])
            take = fp[1] < fr ? 1 : 0;
        else if (fr < fv[2])
            take = 0;
        else if (fr < fv[3])
            take = fp[2] <= fr ? 2 : -1;
        else
            take = fp[3] < fv[3] ? 3 : -1;
        if (take >= 0) {
            for (int k = 0; k < 3; k++) v[3][k] = pts[take][k];
            fv[3] = fp[take];
            continue;
        }
        for (int i = 1; i < 4; i++)  // shrink towards the best vertex
            for (int k = 0; k < 3; k++) v[i][k] = v[0][k] + 0.5 * (v[i][k] - v[0][k]);
        SRED_TRY(e.run(v + 1, 3, fv + 1));
    }
    int b = 0;
    for (int i = 1; i < 4; i++)
        if (fv[i] < fv[b]) b = i;
    for (int k = 0; k < 3; k++) params[k] = v[b][k];
    if (error) *error = fv[b];
    ctx->last_launches = (int)std::min<int64_t>(e.launches, 1 << 30);
    ctx->last_pairs = e.pixel_evals;
    ctx->summed_ms = (float)e.rows_ms;
    return SRED_OK;
}

}  // namespace

}  // namespace sred
