// quadtree.hip -- the Quad/Octree estimate of the GAT parameters (the
// "Quad/Octree" branch of Utils.getInputImage2D, Utils.java:686-703, over
// QuadTree_.java:135-427).
//
// Reference algorithm (QuadTree_(w, h, 4, 0.01f), robust means with 50
// iterations, LTS variances with fraction 0.75f):
//   * buildTree: a node divides into four (w/2 x h/2 each; an odd side drops
//     its last column or row) unless w <= minSize, h <= minSize or the
//     Fisher test finds its variance indistinguishable from its noise
//     variance: min/max of (block variance, pseudo-residual variance) <
//     F^-1(alpha; n-1, n-1), both variances as float sums in raster order.
//   * per leaf a Leclerc-weighted robust mean and the least-trimmed-squares
//     variance (the floor(0.75 n) smallest squared residuals, sorted, summed
//     in float);
//   * a least-squares line of variance against mean over the leaves in
//     pre-order (commons-math3 SimpleRegression): g0 = slope, eDC =
//     intercept; the transform (2/g0) sqrt(g0 v + 3/8 g0^2 + eDC).
//
// MI355X mapping.  Every node of one level has the same size W_k x H_k, so
// the nodes of level k form a complete 2^k x 2^k grid of cells (i, j) whose
// origins are x = sum over levels l = 1..k of bit (k - l) of i times W_l
// (a child sits at its parent's origin plus the parent's half size, which
// for odd sides is not i W_k), and their criteria do not depend on the
// decisions above them: one launch evaluates every node that could divide
// (k_qt_criteria, one wave per node, the root first since its chain is the
// longest); the host then reads the 8 bytes per node, applies the Fisher
// thresholds (one F quantile per level) and walks the tree in pre-order.
// The sums keep the reference's sequential float order: the wave's lanes
// load and transform 64 consecutive elements at once (coalesced) and park
// the terms in LDS; then lane c runs the in-order chain of term stream c
// (for a node's first pass: sum v, sum v*v and sum of residuals on lanes 0,
// 1, 2 in the same instructions), reading LDS 4 elements at a time, while
// the next 512 elements' loads are in flight.  The robust means
// (k_qt_robust_mean) and LTS sums (k_qt_lts_sum) use the same chain per
// leaf; the LTS sort is a segmented radix sort (hipCUB) of the leaves'
// squared residuals.  Results equal the CPU restatement bit for bit except
// where the double exp of the robust-mean weights differs by an ulp from the
// host's and changes a float (probability ~1e-9 per weight).
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "sred_internal.h"
#include "chain.h"

namespace sred {

namespace {

constexpr int kQtMaxLevels = 16;

struct QtLevels {
    int nlev;  // levels whose nodes get a criterion (0 .. nlev-1)
    int W[kQtMaxLevels], H[kQtMaxLevels];
    long long first[kQtMaxLevels + 1];  // first node index of each level
};

// Pixel offset of element e (raster order) of a nw-wide node at (x0, y0).
__device__ __forceinline__ int64_t node_pixel(int e, int nw, int x0, int y0, int w)
{
    const int yy = (int)((unsigned)e / (unsigned)nw), xx = e - yy * nw;
    return (int64_t)(y0 + yy) * w + x0 + xx;
}

// Pseudo-residuals (5 z - neighbours inside the image) * scale for every
// pixel (QuadTree_.java:199-212), and a flag for a non-finite pixel
// (Utils.getMeanAndVarBlock2D on the root, Utils.java:879-883).
__global__ __launch_bounds__(256) void k_qt_residuals(const float *__restrict__ img, int w, int h, float scale,
                                                      float *__restrict__ res, int *__restrict__ bad)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * blockDim.x + threadIdx.x, y = blockIdx.y;
    if (x >= w) return;
    const float *p = img + (int64_t)y * w + x;
    const float z = p[0];
    if (!(fabsf(z) <= 3.402823466e38f)) atomicOr(bad, 1);
    float lap = 5.0f * z;
    if (x > 0) lap -= p[-1];
    if (x < w - 1) lap -= p[1];
    if (y > 0) lap -= p[-w];
    if (y < h - 1) lap -= p[w];
    res[(int64_t)y * w + x] = lap * scale;
}

// stats[node] = (block variance, noise variance) of every node of levels
// 0 .. nlev-1 (QuadTree_.computeCriterion, :170-232); one wave per node.
__global__ __launch_bounds__(64) void k_qt_criteria(const float *__restrict__ img, const float *__restrict__ res, int w,
                                                    QtLevels L, float2 *__restrict__ stats)
{
#pragma clang fp contract(off)
    __shared__ __align__(16) float sb[3 * kStride + kPad];
    const long long node = blockIdx.x;
    int k = 0;
    while (k + 1 < L.nlev && node >= L.first[k + 1]) k++;
    const long long idx = node - L.first[k];
    const int i = (int)(idx & ((1LL << k) - 1)), j = (int)(idx >> k);
    const int nw = L.W[k], nh = L.H[k], n = nw * nh;
    int x0 = 0, y0 = 0;
    for (int l = 1; l <= k; l++) {
        x0 += ((i >> (k - l)) & 1) * L.W[l];
        y0 += ((j >> (k - l)) & 1) * L.H[l];
    }

    // pass 1: sum v, sum v*v (Utils.java:876-887) and the sum of residuals
    // (:219-223): three in-order chains on lanes 0, 1, 2
    float acc = 0.0f;
    auto fetch1 = [&](int e) {
        float2 r = make_float2(0.0f, 0.0f);
        if (e >= 0) {
            const int64_t p = node_pixel(e, nw, x0, y0, w);
            r = make_float2(img[p], res[p]);
        }
        return r;
    };
    auto term1 = [](float2 r, float (&t)[3]) {
        t[0] = r.x;
        t[1] = r.x * r.x;
        t[2] = r.y;
    };
    auto add = [&](int, float x) { acc += x; };
    stream_chain<3>(n, fetch1, term1, add, sb);
    const float fn = (float)n;
    const float mean = __shfl(acc, 0) / fn;
    const float bvar = __shfl(acc, 1) / fn - mean * mean;
    const float mr = __shfl(acc, 2) / fn;

    // pass 2: sum of (float)((r - mean_r)^2), the square in double (:226-229)
    acc = 0.0f;
    auto fetch2 = [&](int e) { return e >= 0 ? res[node_pixel(e, nw, x0, y0, w)] : 0.0f; };
    auto term2 = [&](float r, float (&t)[1]) {
        const double d = (double)r - (double)mr;
        t[0] = (float)(d * d);
    };
    stream_chain<1>(n, fetch2, term2, add, sb);
    if (threadIdx.x == 0) stats[node] = make_float2(bvar, acc / (fn - 1.0f));
}

// QuadTree_.calculateRobustMean (:264-312) per leaf; one wave per leaf.
__global__ __launch_bounds__(64) void k_qt_robust_mean(const float *__restrict__ img, int w,
                                                       const int4 *__restrict__ leaves, int max_iter,
                                                       float *__restrict__ means)
{
#pragma clang fp contract(off)
    __shared__ __align__(16) float sb[2 * kStride + kPad];
    const int4 lf = leaves[blockIdx.x];
    const int x0 = lf.x, y0 = lf.y, nw = lf.z, n = lf.z * lf.w;
    auto fetch = [&](int e) { return e >= 0 ? img[node_pixel(e, nw, x0, y0, w)] : 0.0f; };

    // one-pass estimate on lane 0: the variance term uses the running sum (as written)
    float mean = 0.0f, var = 0.0f;
    auto term0 = [](float v, float (&t)[1]) { t[0] = v; };
    auto op0 = [&](int, float v) {
        mean += v;
        const float d = v - mean;
        var += d * d;
    };
    stream_chain<1>(n, fetch, term0, op0, sb);
    mean = __shfl(mean, 0);
    var = __shfl(var, 0);
    mean = mean / (float)n;
    var = var / (float)(n - 1);
    var = java_maxf(var, kEps);
    const double two_var = 2.0 * (double)var;

    float prev;
    int it = 0;
    do {
        prev = mean;
        float acc = 0.0f;  // lane 0: weight sum, lane 1: weighted sum
        auto term = [&](float v, float (&t)[2]) {
            const double d = (double)(v - mean);
            const float wt = (float)exp(-(d * d) / two_var);
            t[0] = wt;
            t[1] = wt * v;
        };
        auto add = [&](int, float x) { acc += x; };
        stream_chain<2>(n, fetch, term, add, sb);
        mean = __shfl(acc, 1) / java_maxf(__shfl(acc, 0), kEps);
        it++;
    } while ((double)fabsf(mean - prev) > 1e-5 && it < max_iter);
    if (threadIdx.x == 0) means[blockIdx.x] = mean;
}

// Squared residuals about the robust mean, leaf by leaf into seg[off[leaf] ..].
__global__ __launch_bounds__(64) void k_qt_lts_residuals(const float *__restrict__ img, int w,
                                                         const int4 *__restrict__ leaves, const int *__restrict__ off,
                                                         const float *__restrict__ means, float *__restrict__ seg)
{
#pragma clang fp contract(off)
    const int4 lf = leaves[blockIdx.x];
    const int n = lf.z * lf.w;
    const float m = means[blockIdx.x];
    float *out = seg + off[blockIdx.x];
    for (int e = threadIdx.x; e < n; e += 64) {
        const float d = img[node_pixel(e, lf.z, lf.x, lf.y, w)] - m;
        out[e] = d * d;
    }
}

// ---- host side ----

// Regularised incomplete beta I_x(a, b) = x^a (1-x)^b / (a B(a, b)) / g, g
// the continued fraction 1 + d1/(1 + d2/(1 + ...)) of Abramowitz & Stegun
// 26.5.8 by Lentz's method; reflected where it converges slowly.
double beta_reg(double a, double b, double x)
{
    if (x <= 0.0) return 0.0;
    if (x >= 1.0) return 1.0;
    const bool flip = x >= (a + 1.0) / (a + b + 2.0);
    if (flip) {
        std::swap(a, b);
        x = 1.0 - x;
    }
    const double tiny = 1e-300;
    double g = 1.0, C = 1.0, D = 0.0;
    for (long m = 0; m < 2000000; m++) {
        const double dodd = -(a + m) * (a + b + m) * x / ((a + 2.0 * m) * (a + 2.0 * m + 1.0));
        const double deven = (m + 1.0) * (b - m - 1.0) * x / ((a + 2.0 * m + 1.0) * (a + 2.0 * m + 2.0));
        double delta = 1.0;
        for (double dm : {dodd, deven}) {
            D = 1.0 + dm * D;
            if (std::fabs(D) < tiny) D = tiny;
            C = 1.0 + dm / C;
            if (std::fabs(C) < tiny) C = tiny;
            D = 1.0 / D;
            delta = C * D;
            g *= delta;
        }
        if (std::fabs(delta - 1.0) < 1e-16) break;
    }
    const double lfront = std::lgamma(a + b) - std::lgamma(a) - std::lgamma(b) + a * std::log(x) + b * std::log1p(-x);
    const double v = std::exp(lfront) / (a * g);
    return flip ? 1.0 - v : v;
}

// SimpleRegression with intercept (commons-math3 3.6.1): updating sums, then
// slope = sumXY/sumXX and intercept = (sumY - slope sumX)/n.
void simple_regression(const double *x, const double *y, int64_t n, double *slope, double *intercept)
{
#pragma clang fp contract(off)
    double sx = 0.0, sy = 0.0, sxx = 0.0, sxy = 0.0, xbar = 0.0, ybar = 0.0;
    for (int64_t i = 0; i < n; i++) {
        if (i == 0) {
            xbar = x[0];
            ybar = y[0];
        } else {
            const double fact1 = 1.0 + (double)i, fact2 = (double)i / (1.0 + (double)i);
            const double dx = x[i] - xbar, dy = y[i] - ybar;
            sxx += dx * dx * fact2;
            sxy += dx * dy * fact2;
            xbar += dx / fact1;
            ybar += dy / fact1;
        }
        sx += x[i];
        sy += y[i];
    }
    double s = NAN;
    if (n >= 2 && !(std::fabs(sxx) < 10.0 * 4.9406564584124654e-324)) s = sxy / sxx;
    *slope = s;
    *intercept = (sy - s * sx) / (double)n;
}

}  // namespace

// Inverse CDF of F(d1, d2) by bisection on I_{d1 x/(d1 x + d2)}(d1/2, d2/2)
// to adjacent doubles (FDistribution.inverseCumulativeProbability).
double f_quantile(double d1, double d2, double p)
{
    if (!(p > 0.0)) return 0.0;
    if (!(p < 1.0)) return INFINITY;
    auto cdf = [&](double x) { return beta_reg(0.5 * d1, 0.5 * d2, d1 * x / (d1 * x + d2)); };
    double lo = 0.0, hi = 1.0;
    while (cdf(hi) < p) {
        lo = hi;
        hi *= 2.0;
    }
    for (int it = 0; it < 2100; it++) {
        const double mid = 0.5 * (lo + hi);
        if (mid <= lo || mid >= hi) break;
        (cdf(mid) < p ? lo : hi) = mid;
    }
    return 0.5 * (lo + hi);
}

void simple_regression_host(const double *x, const double *y, int64_t n, double *slope, double *intercept)
{
    simple_regression(x, y, n, slope, intercept);
}

int quadtree_gat_dev(sred_ctx *ctx, const float *d_img, int w, int h, int min_size, float alpha, int max_iter,
                     float frac, double *g0_edc, int *n_leaves, int cap, int32_t *rects, float *means, float *vars)
{
    if (w < 1 || h < 1 || (int64_t)w * h >= (1LL << 31) || min_size < 1 || !(alpha >= 0.0f && alpha <= 1.0f) ||
        !(frac >= 0.0f && frac <= 1.0f) || cap < 0) {
        set_error("quadtree GAT: needs w, h >= 1 (w*h < 2^31), min_size >= 1, alpha and the LTS fraction in [0, 1]");
        return SRED_ERR_ARG;
    }
    if ((int64_t)w * h == 1) {
        set_error("quadtree GAT: a one-pixel image has no F test (FDistribution(0, 0))");
        return SRED_ERR_ARG;
    }
    // level geometry: nodes of level k are W[k] x H[k]; level k+1 exists while k's can divide
    QtLevels L{};
    int W = w, H = h, D = 0;
    std::vector<int> LW{W}, LH{H};
    while (W > min_size && H > min_size) {
        W /= 2;
        H /= 2;
        LW.push_back(W);
        LH.push_back(H);
        D++;
    }
    L.nlev = std::max(D, 1);  // the root's criterion is evaluated even when it cannot divide
    if (L.nlev > kQtMaxLevels - 1) {
        set_error("quadtree GAT: too many levels");
        return SRED_ERR_ARG;
    }
    long long nodes = 0;
    for (int k = 0; k < L.nlev; k++) {
        L.W[k] = LW[k];
        L.H[k] = LH[k];
        L.first[k] = nodes;
        nodes += 1LL << (2 * k);
    }
    L.first[L.nlev] = nodes;
    if (nodes > (1LL << 27)) {
        set_error("quadtree GAT: too many tree nodes (raise min_size)");
        return SRED_ERR_ARG;
    }
    const int64_t npx = (int64_t)w * h;
    SRED_TRY(ctx->qt_res.ensure(sizeof(float) * npx + 256));
    SRED_TRY(ctx->qt_nodes.ensure(sizeof(float2) * nodes + 256));
    SRED_TRY(ctx->misc.ensure(kMiscBytes));
    int *d_bad = reinterpret_cast<int *>(ctx->misc.as<char>() + 2048);
    SRED_TRY(zero_dev(ctx, d_bad, sizeof(int)));
    const float scale = (float)(1.0 / std::sqrt(30.0));  // (float)(1/sqrt(l*l + l)), l = 5 (QuadTree_.java:193)
    hipLaunchKernelGGL(k_qt_residuals, dim3((w + 255) / 256, h), dim3(256), 0, ctx->stream, d_img, w, h, scale,
                       ctx->qt_res.as<float>(), d_bad);
    SRED_TRY(time_begin(ctx));
    hipLaunchKernelGGL(k_qt_criteria, dim3((unsigned)nodes), dim3(64), 0, ctx->stream, d_img, ctx->qt_res.as<float>(),
                       w, L, ctx->qt_nodes.as<float2>());
    SRED_TRY(time_end(ctx, npx * L.nlev));
    SRED_HIP_CHECK(hipGetLastError());
    std::vector<float2> st((size_t)nodes);
    int bad = 0;
    SRED_HIP_CHECK(hipMemcpyAsync(&bad, d_bad, sizeof(int), hipMemcpyDeviceToHost, ctx->stream));
    SRED_HIP_CHECK(hipMemcpyAsync(st.data(), ctx->qt_nodes.ptr, sizeof(float2) * nodes, hipMemcpyDeviceToHost,
                                  ctx->stream));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    if (bad) {
        set_error("Input contains invalid (NaN or Infinite) values.");
        return SRED_ERR_NONFINITE;
    }
    // Fisher thresholds, one per level (QuadTree_.performFisherTest, :236-249)
    std::vector<float> thr(L.nlev);
    for (int k = 0; k < L.nlev; k++) {
        const double dof = (double)L.W[k] * L.H[k] - 1.0;
        thr[k] = (float)f_quantile(dof, dof, (double)alpha);
    }
    // pre-order walk: top-left, top-right, bottom-left, bottom-right; a
    // node carries its origin (a child's is its parent's plus the half size)
    std::vector<int4> leaves;
    struct Item { int k, i, j, x, y; };
    std::vector<Item> stack{{0, 0, 0, 0, 0}};
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        bool divide = false;
        if (it.k < L.nlev) {
            const float2 s = st[(size_t)(L.first[it.k] + ((long long)it.j << it.k) + it.i)];
            if (s.x != s.x) {
                set_error("Computed variance is NaN. Check inputs and computations.");
                return SRED_ERR_NONFINITE;
            }
            const float ratio = java_minf(s.x, s.y) / java_maxf(s.x, s.y);
            divide = it.k < D && ratio < thr[it.k];
        }
        if (!divide) {
            leaves.push_back(make_int4(it.x, it.y, LW[it.k], LH[it.k]));
            continue;
        }
        const int k1 = it.k + 1, i2 = 2 * it.i, j2 = 2 * it.j, hw = LW[k1], hh = LH[k1];
        stack.push_back({k1, i2 + 1, j2 + 1, it.x + hw, it.y + hh});  // pushed in reverse: popped TL, TR, BL, BR
        stack.push_back({k1, i2, j2 + 1, it.x, it.y + hh});
        stack.push_back({k1, i2 + 1, j2, it.x + hw, it.y});
        stack.push_back({k1, i2, j2, it.x, it.y});
    }
    const int nl = (int)leaves.size();
    if (n_leaves) *n_leaves = nl;
    std::vector<int> off(nl + 1, 0);
    for (int i = 0; i < nl; i++) off[i + 1] = off[i] + leaves[i].z * leaves[i].w;
    const int nseg = off[nl];

    SRED_TRY(ctx->qt_leaves.ensure(sizeof(int4) * nl + sizeof(int) * (nl + 1) + 2 * sizeof(float) * nl + 1024));
    int4 *d_leaves = ctx->qt_leaves.as<int4>();
    int *d_off = reinterpret_cast<int *>(d_leaves + nl);
    float *d_means = reinterpret_cast<float *>(d_off + nl + 1), *d_vars = d_means + nl;
    SRED_HIP_CHECK(hipMemcpyAsync(d_leaves, leaves.data(), sizeof(int4) * nl, hipMemcpyHostToDevice, ctx->stream));
    SRED_HIP_CHECK(hipMemcpyAsync(d_off, off.data(), sizeof(int) * (nl + 1), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_qt_robust_mean, dim3(nl), dim3(64), 0, ctx->stream, d_img, w, d_leaves, max_iter, d_means);
    SRED_TRY(ctx->qt_seg.ensure(2 * sizeof(float) * nseg + 256));
    float *seg_in = ctx->qt_seg.as<float>(), *seg_out = seg_in + nseg;
    hipLaunchKernelGGL(k_qt_lts_residuals, dim3(nl), dim3(64), 0, ctx->stream, d_img, w, d_leaves, d_off, d_means,
                       seg_in);
    size_t tmp = 0;
    SRED_HIP_CHECK(hipcub::DeviceSegmentedRadixSort::SortKeys(nullptr, tmp, seg_in, seg_out, nseg, nl, d_off,
                                                              d_off + 1, 0, 32, ctx->stream));
    SRED_TRY(ctx->qt_tmp.ensure(tmp + 256));
    SRED_HIP_CHECK(hipcub::DeviceSegmentedRadixSort::SortKeys(ctx->qt_tmp.ptr, tmp, seg_in, seg_out, nseg, nl, d_off,
                                                              d_off + 1, 0, 32, ctx->stream));
    hipLaunchKernelGGL(k_qt_lts_sum, dim3(nl), dim3(64), 0, ctx->stream, seg_out, d_off, frac, d_vars);
    SRED_HIP_CHECK(hipGetLastError());
    std::vector<float> mv(2 * (size_t)nl);
    SRED_HIP_CHECK(hipMemcpyAsync(mv.data(), d_means, sizeof(float) * 2 * nl, hipMemcpyDeviceToHost, ctx->stream));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    const int nc = std::min(nl, cap);
    for (int i = 0; i < nc; i++) {
        if (rects) {
            rects[4 * i] = leaves[i].x;
            rects[4 * i + 1] = leaves[i].y;
            rects[4 * i + 2] = leaves[i].z;
            rects[4 * i + 3] = leaves[i].w;
        }
        if (means) means[i] = mv[i];
        if (vars) vars[i] = mv[nl + i];
    }
    if (nl < 3) {
        set_error("quadtree GAT: fewer than 3 leaves for the regression (SimpleRegression.regress needs 3 points)");
        return SRED_ERR_ARG;
    }
    std::vector<double> x(nl), y(nl);
    for (int i = 0; i < nl; i++) {
        x[i] = (double)mv[i];
        y[i] = (double)mv[nl + i];
    }
    simple_regression(x.data(), y.data(), nl, &g0_edc[0], &g0_edc[1]);
    return SRED_OK;
}

int gat_tree_dev(sred_ctx *ctx, float *d_img, int64_t n, double g0, double edc)
{
#pragma clang fp contract(off)
    // QuadTree_.applyGATtree (:409-425): ref = 3/8 g0^2 + eDC, then the GAT kernel
    return gat_ref_dev(ctx, d_img, n, g0, (3.0 / 8.0) * g0 * g0 + edc);
}

}  // namespace sred
