// hu.hip -- Hu template metric: the first Hu invariant of every window
// against that of one template.
//
// Reference: kernelGetPatchHu2D.cl:11-282 (an OpenCL kernel with no Java
// caller).  Per interior pixel, in FP64:
//   * window mean over the square minus its four corners (the kernel's
//     predicate uses integer division, :129-139) divided by K (:140);
//   * mean subtraction, min/max and min-max normalisation over the float
//     ellipse (:143-182);
//   * raw moments with 1-based coordinates (i+1, j+1), central moments with
//     0-based ones (:190-234); hu1 = nu20 + nu02, nu = mu / (M00^2 + eps);
//   * out = (float) sqrt(hu1_template * hu1_window) (:267).
// The template's hu1 is the same for every pixel (the kernel recomputes it
// per work-item, :36-101) and is computed once on the host here, with the
// kernel's predicates (float for the copy, double for the moments).
// pow(x, 2.0) is x*x, the correctly rounded square.  No FMA contraction, so
// the sums are the kernel's sequential FP64 sums.
//
// Layout: one workgroup per 16x16 output tile; the input tile with its
// block-radius halo is staged in LDS as float once; every thread walks two
// offset lists (square-minus-corners for the mean, float ellipse for the
// rest) in the kernel's row-major order.  FP64 VALU bound (~25 FP64 ops per
// ellipse pixel including two divisions), 8 bytes of HBM per pixel.
#include <float.h>

#include <algorithm>
#include <cmath>
#include <vector>

#include "sred_internal.h"

namespace sred {

constexpr int kHuTile = 16;

// float ellipse predicate of the kernel (e.g. kernelGetPatchHu2D.cl:38-40)
static bool hu_in_f(int i, int j, int brw, int brh)
{
#pragma clang fp contract(off)
    float dx = (float)(i - brw) / (float)brw;
    float dy = (float)(j - brh) / (float)brh;
    float a = dx * dx;
    float b = dy * dy;
    return a + b <= 1.0f;
}
// double ellipse predicate (kernelGetPatchHu2D.cl:58-60)
static bool hu_in_d(int i, int j, int brw, int brh)
{
#pragma clang fp contract(off)
    double dx = (double)(i - brw) / brw;
    double dy = (double)(j - brh) / brh;
    double a = dx * dx;
    double b = dy * dy;
    return a + b <= 1.0;
}

// hu1 of the template, kernelGetPatchHu2D.cl:36-101
double hu1_template_host(const float *tmpl, int bw, int bh, double eps)
{
#pragma clang fp contract(off)
    const int brw = bw / 2, brh = bh / 2;
    std::vector<double> p((size_t)bw * bh, 0.0);
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++)
            if (hu_in_f(i, j, brw, brh)) p[j * bw + i] = (double)tmpl[j * bw + i];
    double M00 = 0.0, M10 = 0.0, M01 = 0.0;
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++)
            if (hu_in_d(i, j, brw, brh)) {
                const double v = p[j * bw + i];
                M00 += v;
                M10 += (double)(i + 1) * v;
                M01 += (double)(j + 1) * v;
            }
    const double cx = M10 / (M00 + eps), cy = M01 / (M00 + eps);
    double mu20 = 0.0, mu02 = 0.0;
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++)
            if (hu_in_d(i, j, brw, brh)) {
                const double v = p[j * bw + i];
                const double ddx = (double)i - cx, ddy = (double)j - cy;
                mu20 += (ddx * ddx) * v;
                mu02 += (ddy * ddy) * v;
            }
    const double m2 = M00 * M00;
    return mu20 / (m2 + eps) + mu02 / (m2 + eps);
}

// Per pixel of a (16+2brw) x (16+2brh) LDS tile.  msq: offsets (in tile
// pixels) of the mean set; ell: offsets of the float-ellipse cells; ij: their
// (i, j) patch coordinates packed as i | j << 16.
__global__ __launch_bounds__(256) void k_block_hu(const float *__restrict__ img, int w, int h, int brw, int brh,
                                                  const int *__restrict__ msq, int nsq,
                                                  const int *__restrict__ ell, const int *__restrict__ ij, int K,
                                                  double hu_ref, double eps, float *__restrict__ out)
{
#pragma clang fp contract(off)
    extern __shared__ float tile[];
    const int tw = kHuTile + 2 * brw, th = kHuTile + 2 * brh;
    const int x0 = blockIdx.x * kHuTile, y0 = blockIdx.y * kHuTile;
    for (int i = threadIdx.x; i < tw * th; i += blockDim.x) {
        const int ly = i / tw, lx = i - ly * tw;
        const int gy = y0 - brh + ly, gx = x0 - brw + lx;
        tile[i] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? img[(int64_t)gy * w + gx] : 0.0f;
    }
    __syncthreads();
    const int tx = threadIdx.x % kHuTile, ty = threadIdx.x / kHuTile;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx >= w || gy >= h) return;
    const int64_t o = (int64_t)gy * w + gx;
    if (gx < brw || gx >= w - brw || gy < brh || gy >= h - brh) {
        out[o] = 0.0f;
        return;
    }
    const float *c = tile + ty * tw + tx;  // top-left corner of this pixel's window
    double mean = 0.0;
    for (int k = 0; k < nsq; k++) mean += (double)c[msq[k]];
    mean /= (double)K;
    double mn = DBL_MAX, mx = -DBL_MAX;
    for (int k = 0; k < K; k++) {
        const double v = (double)c[ell[k]] - mean;
        mn = fmin(mn, v);
        mx = fmax(mx, v);
    }
    const double range = mx - mn + eps;
    double M00 = 0.0, M10 = 0.0, M01 = 0.0;
    for (int k = 0; k < K; k++) {
        const double v = ((double)c[ell[k]] - mean - mn) / range;
        const int q = ij[k];
        M00 += v;
        M10 += (double)((q & 0xffff) + 1) * v;
        M01 += (double)((q >> 16) + 1) * v;
    }
    const double cx = M10 / (M00 + eps), cy = M01 / (M00 + eps);
    double mu20 = 0.0, mu02 = 0.0;
    for (int k = 0; k < K; k++) {
        const double v = ((double)c[ell[k]] - mean - mn) / range;
        const int q = ij[k];
        const double ddx = (double)(q & 0xffff) - cx, ddy = (double)(q >> 16) - cy;
        mu20 += (ddx * ddx) * v;
        mu02 += (ddy * ddy) * v;
    }
    const double m2 = M00 * M00;
    const double hu = mu20 / (m2 + eps) + mu02 / (m2 + eps);
    out[o] = (float)sqrt(hu_ref * hu);
}

int block_hu_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, const float *h_tmpl, float *d_out)
{
    const double eps = java_float_literal(kEps);
    const double hu_ref = hu1_template_host(h_tmpl, g.bw, g.bh, eps);
    const int tw = kHuTile + 2 * g.brw, th = kHuTile + 2 * g.brh;
    // offset lists in the kernel's row-major order, relative to the window's top-left corner in the tile
    std::vector<int> lists;
    int nsq = 0, nell = 0;
    for (int j = 0; j < g.bh; j++)
        for (int i = 0; i < g.bw; i++) {
            const int di = (i - g.brw) / g.brw, dj = (j - g.brh) / g.brh;  // integer division, :131-132
            if (!(di != 0 && dj != 0)) {
                lists.push_back(j * tw + i);
                nsq++;
            }
        }
    std::vector<int> ell, ij;
    for (int j = 0; j < g.bh; j++)
        for (int i = 0; i < g.bw; i++)
            if (hu_in_f(i, j, g.brw, g.brh)) {
                ell.push_back(j * tw + i);
                ij.push_back(i | (j << 16));
                nell++;
            }
    if (nell != g.K) {
        set_error("Hu footprint does not match the block size");
        return SRED_ERR_ARG;
    }
    lists.insert(lists.end(), ell.begin(), ell.end());
    lists.insert(lists.end(), ij.begin(), ij.end());
    SRED_TRY(ctx->offs.ensure(sizeof(int) * lists.size()));
    ctx->offs_key[0] = -1;  // the footprint offset cache no longer holds these
    SRED_HIP_CHECK(hipMemcpyAsync(ctx->offs.ptr, lists.data(), sizeof(int) * lists.size(), hipMemcpyHostToDevice,
                                  ctx->stream));
    const int *d_l = ctx->offs.as<int>();
    const size_t lds = sizeof(float) * (size_t)tw * th;
    dim3 grid((g.w + kHuTile - 1) / kHuTile, (g.h + kHuTile - 1) / kHuTile);
    SRED_TRY(time_begin(ctx));
    hipLaunchKernelGGL(k_block_hu, grid, dim3(kHuTile * kHuTile), lds, ctx->stream, d_img, g.w, g.h, g.brw, g.brh,
                       d_l, nsq, d_l + nsq, d_l + nsq + nell, g.K, hu_ref, eps, d_out);
    SRED_HIP_CHECK(hipGetLastError());
    SRED_TRY(time_end(ctx, g.n()));
    // the host list must outlive the asynchronous copy
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SRED_OK;
}

// Template preparation for the Hu map: the window steps of
// kernelGetPatchHu2D.cl:127-182 applied to the template rectangle.
int hu_prepare_template_host(const float *blk, int bw, int bh, int K, float *out)
{
#pragma clang fp contract(off)
    const int brw = bw / 2, brh = bh / 2;
    const double eps = java_float_literal(kEps);
    double mean = 0.0;
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++) {
            const int di = (i - brw) / brw, dj = (j - brh) / brh;
            if (!(di != 0 && dj != 0)) mean += (double)blk[j * bw + i];
        }
    mean /= (double)K;
    double mn = DBL_MAX, mx = -DBL_MAX;
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++)
            if (hu_in_f(i, j, brw, brh)) {
                const double v = (double)blk[j * bw + i] - mean;
                mn = std::min(mn, v);
                mx = std::max(mx, v);
            }
    for (int j = 0; j < bh; j++)
        for (int i = 0; i < bw; i++)
            out[j * bw + i] =
                hu_in_f(i, j, brw, brh) ? (float)(((double)blk[j * bw + i] - mean - mn) / (mx - mn + eps)) : 0.0f;
    return SRED_OK;
}

}  // namespace sred
