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
//     k_gat_rows takes one (parameter set, image row, block) per thread: the
//     sequential double sums of v and of the float squares v*v over the row's
//     pixels of the block; k_gat_error adds the row sums of each block in row
//     order and the block errors in raster order, so an evaluation is
//     deterministic and equals the CPU restatement bit for bit (IEEE double
//     arithmetic and correctly rounded square roots on both).
//   * The minimiser: ImageJ's Minimizer (restarts, random initial simplex)
//     is third-party and not restated; this is a Nelder-Mead simplex with
//     the Lagarias et al. (1998) coefficients, initial simplex x0 + d_i e_i
//     with d = (1, 10, 100), stopping after max_iter iterations or when all
//     vertices have the same objective (the reference's setMaxError(0)).
//     Each iteration evaluates the reflection, expansion and both
//     contractions in one launch (four parameter sets) and then takes the
//     branch the values select, so the path is the sequential algorithm's.
#include <cmath>
#include <vector>

#include "sred_internal.h"

namespace sred {

constexpr int kGatSets = 4;  // parameter sets per objective launch

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

// rows[set][j][bx] = (sum v, sum v*v) over the row j of block column bx
__global__ __launch_bounds__(256) void k_gat_rows(const float *__restrict__ px, int w, int bw, int nbx, int nrows,
                                                  int nsets, GatSets ps, double2 *__restrict__ rows)
{
#pragma clang fp contract(off)
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nsets * nrows * nbx) return;
    const int set = t / (nrows * nbx), rem = t - set * (nrows * nbx);
    const int j = rem / nbx, bx = rem - j * nbx;
    const double gain = ps.gain[set], ref = ps.ref[set];
    const double lim = -ref / gain;
    const float *p = px + (int64_t)j * w + (int64_t)bx * bw;
    double rm = 0.0, rs = 0.0;
    for (int i = 0; i < bw; i++) {
        double d = (double)p[i];
        if (d <= lim)
            d = 0.0;
        else
            d = (2.0 / gain) * sqrt(gain * d + ref);
        const float v = (float)d;
        rm += v;
        rs += v * v;
    }
    rows[t] = make_double2(rm, rs);
}

// err[set] from the row sums: blocks in raster order, rows of a block in order
__global__ __launch_bounds__(128) void k_gat_error(const double2 *__restrict__ rows, int bw, int bh, int nbx, int nby,
                                                   int nrows, double *__restrict__ err)
{
#pragma clang fp contract(off)
    __shared__ double be[128];
    const int set = blockIdx.x, b = threadIdx.x, nb = nbx * nby;
    if (b < nb) {
        const int by = b / nbx, bx = b - by * nbx;
        const double2 *r = rows + (size_t)set * nrows * nbx;
        double mean = 0.0, sq = 0.0;
        for (int j = by * bh; j < (by + 1) * bh; j++) {
            mean += r[(size_t)j * nbx + bx].x;
            sq += r[(size_t)j * nbx + bx].y;
        }
        const int n = bw * bh;
        mean = mean / n;
        const double var = sq / n - mean * mean;
        const double delta = var - 1.0;
        be[b] = (delta * delta) / (double)nb;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        double e = 0.0;
        for (int i = 0; i < nb; i++) e += be[i];
        err[set] = e;
    }
}

static double gat_ref(double gain, double sigma, double offset)
{
#pragma clang fp contract(off)
    return (3.0 / 8.0) * gain * gain + sigma * sigma - gain * offset;
}

int gat_dev(sred_ctx *ctx, float *d_img, int64_t n, double gain, double sigma, double offset)
{
    if (n <= 0) return SRED_OK;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 16384);
    hipLaunchKernelGGL(k_gat, dim3(blocks), dim3(256), 0, ctx->stream, d_img, n, gain, gat_ref(gain, sigma, offset));
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

namespace {

// One device evaluation of up to kGatSets parameter sets (synchronises).
struct GatEval {
    sred_ctx *ctx;
    const float *d_img;
    int w, h, bw, bh, nbx, nby, nrows;
    double2 *d_rows;
    double *d_err;
    int64_t launches = 0;

    int run(const double (*x)[3], int nsets, double *f)
    {
        GatSets ps{};
        for (int s = 0; s < nsets; s++) {
            ps.gain[s] = x[s][0];
            ps.ref[s] = gat_ref(x[s][0], x[s][1], x[s][2]);
        }
        const int threads = nsets * nrows * nbx;
        hipLaunchKernelGGL(k_gat_rows, dim3((threads + 255) / 256), dim3(256), 0, ctx->stream, d_img, w, bw, nbx,
                           nrows, nsets, ps, d_rows);
        hipLaunchKernelGGL(k_gat_error, dim3(nsets), dim3(128), 0, ctx->stream, d_rows, bw, bh, nbx, nby, nrows,
                           d_err);
        SRED_HIP_CHECK(hipGetLastError());
        SRED_HIP_CHECK(hipMemcpyAsync(f, d_err, sizeof(double) * nsets, hipMemcpyDeviceToHost, ctx->stream));
        SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        launches++;
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
    if (e->nbx * e->nby > 128) {
        set_error("unsupported GAT block grid");
        return SRED_ERR_ARG;
    }
    SRED_TRY(ctx->misc.ensure(kMiscBytes));
    SRED_TRY(ctx->pm_part.ensure(sizeof(double2) * (size_t)kGatSets * e->nrows * e->nbx + 256));
    e->d_rows = ctx->pm_part.as<double2>();
    e->d_err = reinterpret_cast<double *>(ctx->misc.as<char>() + 2048);  // 4 doubles in the mask-check area
    return SRED_OK;
}

}  // namespace

int gat_objective_dev(sred_ctx *ctx, const float *d_img, int w, int h, double gain, double sigma, double offset,
                      double *error)
{
    GatEval e;
    SRED_TRY(gat_setup(ctx, d_img, w, h, &e));
    const double x[1][3] = {{gain, sigma, offset}};
    return e.run(x, 1, error);
}

// Nelder-Mead as the CPU restatement (same expressions, no contraction), with
// the candidate points of an iteration evaluated together.
int gat_minimize_dev(sred_ctx *ctx, const float *d_img, int w, int h, double gain0, double sigma0, double offset0,
                     int max_iter, double *params, double *error)
{
#pragma clang fp contract(off)
    GatEval e;
    SRED_TRY(gat_setup(ctx, d_img, w, h, &e));
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
    return SRED_OK;
}

}  // namespace sred
