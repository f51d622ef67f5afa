// octree.hip -- the octree GAT estimate of a volume: OcTree_ (OcTree_.java:
// 136-486) as called by the "Quad/Octree" branch of Utils.getInputImage3D
// (Utils.java:1234-1254) and by the OcTree_ plugin (:82-96), and its
// transform applyGATtree (:471-485).
//
// The same design as the quadtree (quadtree.hip): the nodes of level k form
// a complete 2^k x 2^k x 2^k grid of boxes W_k x H_k x D_k whose origins are
// sums of the halved sizes picked by the bits of the cell index, their
// criteria do not depend on the decisions above them, so one launch
// evaluates every node that could divide (two waves per node, in-order
// chains: oc_pass1 with its double terms staged by all lanes on one,
// chain.h's float chains on the other), and the host applies the Fisher thresholds and walks
// the tree in pre-order (children x fastest, then y, then z).  What differs
// from the quadtree, as the reference writes it:
//   * the node variance is Utils.getMeanAndVarBlock3D's: double sums of v
//     and v*v (Utils.java:1356-1380), cast to float (:219);
//   * the pseudo-residuals are 7 v minus the six neighbours inside the
//     volume, over sqrt(56) (:230-280);
//   * the F test keeps the quadtree's region size width*height (:225);
//   * a node is a leaf when any side is <= min_size (:178);
//   * the regression's slope and intercept are cast to float (:458-469),
//     and applyGATtree zeroes v < -ref/g0 (:477) where the quadtree tests
//     g0 v + ref < 0;
//   * no non-finite check (getMeanAndVarBlock3D has none).
// getInputImage3D passes width*height as the voxel count to applyGATtree
// (:1253), so it returns one slice and the volume step that follows
// overruns it: the transform covers all voxels, as the OcTree_ plugin's
// (:109) does.
#include <hipcub/hipcub.hpp>

#include <algorithm>
#include <cmath>
#include <vector>

#include "sred_internal.h"
#include "chain.h"

namespace sred {

namespace {

constexpr int kOcMaxLevels = 12;

struct OcLevels {
    int nlev;  // levels whose nodes get a criterion (0 .. nlev-1)
    int W[kOcMaxLevels], H[kOcMaxLevels], D[kOcMaxLevels];
    long long first[kOcMaxLevels + 1];  // first node index of each level
};

struct Box {
    int x, y, z, w, h, d;
};

// Voxel offset of element e (z, y, x order) of a nw x nh x nd box at (x0, y0, z0).
__device__ __forceinline__ int64_t box_voxel(int e, int nw, int nh, int x0, int y0, int z0, int w, int64_t plane)
{
    const unsigned u = (unsigned)e, t = u / (unsigned)nw, zz = t / (unsigned)nh;
    const int xx = (int)(u - t * (unsigned)nw), yy = (int)(t - zz * (unsigned)nh);
    return (int64_t)(z0 + (int)zz) * plane + (int64_t)(y0 + yy) * w + x0 + xx;
}

// Pseudo-residuals (7 v - the neighbours inside the volume) * scale
// (OcTree_.java:244-258); 64 x 4 voxels per workgroup.
__global__ __launch_bounds__(256) void k_oc_residuals(const float *__restrict__ img, int w, int h, int d, float scale,
                                                      float *__restrict__ res)
{
#pragma clang fp contract(off)
    const int x = blockIdx.x * 64 + (threadIdx.x & 63), y = blockIdx.y * 4 + (threadIdx.x >> 6), z = blockIdx.z;
    if (x >= w || y >= h) return;
    const int64_t plane = (int64_t)w * h, o = z * plane + (int64_t)y * w + x;
    const float *p = img + o;
    float lap = 7.0f * p[0];
    if (x > 0) lap -= p[-1];
    if (x < w - 1) lap -= p[1];
    if (y > 0) lap -= p[-w];
    if (y < h - 1) lap -= p[w];
    if (z > 0) lap -= p[-plane];
    if (z < d - 1) lap -= p[plane];
    res[o] = lap * scale;
}

constexpr int kDStride = kSeg + 2;  // doubles per double term stream (streams start 4 banks apart)

// Pass 1 of a node's criterion: the double sums of v (lane 0) and v*v
// (lane 1) and the float sum of the residuals (lane 2), each in element
// order.  All 64 lanes form the double terms while staging a block of
// elements in LDS, so a chain step is one double add and one float add
// (every lane of the three does both; the unused sums are discarded); the
// next block's loads are in flight meanwhile.
template <class Fetch>
__device__ __forceinline__ void oc_pass1(int n, Fetch fetch, double *sd, float *sf, double &accd, float &acc)
{
#pragma clang fp contract(off)
    constexpr int Q = kSeg / 64;
    const int lane = threadIdx.x;
    float2 raw[Q];
#pragma unroll
    for (int q = 0; q < Q; q++) {
        const int e = lane + 64 * q;
        raw[q] = fetch(e < n ? e : -1);
    }
    for (int base = 0; base < n; base += kSeg) {
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const double dv = (double)raw[q].x;
            sd[lane + 64 * q] = dv;
            sd[kDStride + lane + 64 * q] = dv * dv;
            sf[lane + 64 * q] = raw[q].y;
        }
        const int nb = base + kSeg;
#pragma unroll
        for (int q = 0; q < Q; q++) {
            const int e = nb + lane + 64 * q;
            raw[q] = fetch(e < n ? e : -1);
        }
        __syncthreads();
        if (lane < 3) {
            const double *d = sd + (lane == 1 ? kDStride : 0);
            const double2 *d2 = reinterpret_cast<const double2 *>(d);
            const float4 *f4 = reinterpret_cast<const float4 *>(sf);
            const int cnt = min(kSeg, n - base), ng = cnt >> 2;
            auto eat4 = [&](const double2 &a, const double2 &b, const float4 &f) {
                accd += a.x;
                acc += f.x;
                accd += a.y;
                acc += f.y;
                accd += b.x;
                acc += f.z;
                accd += b.y;
                acc += f.w;
            };
            if (cnt == kSeg) {
                // full stage: the next 16 elements' reads are issued before the
                // current 16 are consumed (the last prefetch reads the padding)
                double2 DA[8], DB[8];
                float4 FA[4], FB[4];
#pragma unroll
                for (int u = 0; u < 8; u++) DA[u] = d2[u];
#pragma unroll
                for (int u = 0; u < 4; u++) FA[u] = f4[u];
#pragma unroll 1
                for (int e = 0; e < kSeg; e += 32) {
#pragma unroll
                    for (int u = 0; u < 8; u++) DB[u] = d2[(e + 16) / 2 + u];
#pragma unroll
                    for (int u = 0; u < 4; u++) FB[u] = f4[(e + 16) / 4 + u];
#pragma unroll
                    for (int u = 0; u < 4; u++) eat4(DA[2 * u], DA[2 * u + 1], FA[u]);
#pragma unroll
                    for (int u = 0; u < 8; u++) DA[u] = d2[(e + 32) / 2 + u];
#pragma unroll
                    for (int u = 0; u < 4; u++) FA[u] = f4[(e + 32) / 4 + u];
#pragma unroll
                    for (int u = 0; u < 4; u++) eat4(DB[2 * u], DB[2 * u + 1], FB[u]);
                }
            } else {
#pragma unroll 4
                for (int g = 0; g < ng; g++) eat4(d2[2 * g], d2[2 * g + 1], f4[g]);
                for (int t = ng << 2; t < cnt; t++) {
                    accd += d[t];
                    acc += sf[t];
                }
            }
        }
        __syncthreads();
    }
}

// stats[node] = (block variance, noise variance) of every node of levels
// 0 .. nlev-1 (OcTree_.computeCriterion, :216-280); two single-wave
// workgroups per node, root first: role 0 (even blockIdx) the double sums
// of v and v*v, role 1 the residual sum and then the second pass, so the
// root's double chain runs beside its two float chains.
__global__ __launch_bounds__(64) void k_oc_criteria(const float *__restrict__ img, const float *__restrict__ res, int w,
                                                    int h, OcLevels L, float2 *__restrict__ stats)
{
#pragma clang fp contract(off)
    __shared__ __align__(16) float sb[kStride + kPad];
    __shared__ __align__(16) double sd[2 * kDStride + 16];  // + the read-ahead of a full stage
    __shared__ __align__(16) float sf[kSeg + kPad];
    const long long node = blockIdx.x >> 1;
    const int role = blockIdx.x & 1;
    int k = 0;
    while (k + 1 < L.nlev && node >= L.first[k + 1]) k++;
    const long long idx = node - L.first[k];
    const long long m = (1LL << k) - 1;
    const int i = (int)(idx & m), j = (int)((idx >> k) & m), l = (int)(idx >> (2 * k));
    const int nw = L.W[k], nh = L.H[k], nd = L.D[k], n = nw * nh * nd;
    int x0 = 0, y0 = 0, z0 = 0;
    for (int q = 1; q <= k; q++) {
        x0 += ((i >> (k - q)) & 1) * L.W[q];
        y0 += ((j >> (k - q)) & 1) * L.H[q];
        z0 += ((l >> (k - q)) & 1) * L.D[q];
    }
    const int64_t plane = (int64_t)w * h;

    float *out = reinterpret_cast<float *>(stats) + 2 * node + role;
    if (role == 0) {
        // double sums of v and v*v (lanes 0, 1) in z, y, x order (lane 2's
        // float chain runs over zeros)
        double accd = 0.0;
        float acc = 0.0f;
        auto fetch1 = [&](int e) {
            return make_float2(e >= 0 ? img[box_voxel(e, nw, nh, x0, y0, z0, w, plane)] : 0.0f, 0.0f);
        };
        oc_pass1(n, fetch1, sd, sf, accd, acc);
        const double mean = __shfl(accd, 0) / (double)n;
        const float bvar = (float)(__shfl(accd, 1) / (double)n - mean * mean);
        if (threadIdx.x == 0) *out = bvar;
        return;
    }

    // pass 1 of the residuals: their float sum in z, y, x order
    float acc = 0.0f;
    auto fetch2 = [&](int e) { return e >= 0 ? res[box_voxel(e, nw, nh, x0, y0, z0, w, plane)] : 0.0f; };
    auto term1 = [](float r, float (&t)[1]) { t[0] = r; };
    auto add = [&](int, float x) { acc += x; };
    stream_chain<1>(n, fetch2, term1, add, sb);
    const float fn = (float)n;
    const float mr = __shfl(acc, 0) / fn;

    // pass 2: sum of (float)((r - mean_r)^2), the square in double (:272-276)
    acc = 0.0f;
    auto term2 = [&](float r, float (&t)[1]) {
        const double dd = (double)r - (double)mr;
        t[0] = (float)(dd * dd);
    };
    stream_chain<1>(n, fetch2, term2, add, sb);
    if (threadIdx.x == 0) *out = acc / (fn - 1.0f);
}

// OcTree_.calculateRobustMean (:316-366) per leaf; one wave per leaf.
__global__ __launch_bounds__(64) void k_oc_robust_mean(const float *__restrict__ img, int w, int h,
                                                       const Box *__restrict__ leaves, int max_iter,
                                                       float *__restrict__ means)
{
#pragma clang fp contract(off)
    __shared__ __align__(16) float sb[2 * kStride + kPad];
    const Box b = leaves[blockIdx.x];
    const int n = b.w * b.h * b.d;
    const int64_t plane = (int64_t)w * h;
    auto fetch = [&](int e) { return e >= 0 ? img[box_voxel(e, b.w, b.h, b.x, b.y, b.z, w, plane)] : 0.0f; };

    // one-pass estimate on lane 0: the variance term uses the running sum (as written)
    float mean = 0.0f, var = 0.0f;
    auto term0 = [](float v, float (&t)[1]) { t[0] = v; };
    auto op0 = [&](int, float v) {
        mean += v;
        const float dv = v - mean;
        var += dv * dv;
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
            const double dv = (double)(v - mean);
            const float wt = (float)exp(-(dv * dv) / two_var);
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
__global__ __launch_bounds__(64) void k_oc_lts_residuals(const float *__restrict__ img, int w, int h,
                                                         const Box *__restrict__ leaves, const int *__restrict__ off,
                                                         const float *__restrict__ means, float *__restrict__ seg)
{
#pragma clang fp contract(off)
    const Box b = leaves[blockIdx.x];
    const int n = b.w * b.h * b.d;
    const int64_t plane = (int64_t)w * h;
    const float m = means[blockIdx.x];
    float *out = seg + off[blockIdx.x];
    for (int e = threadIdx.x; e < n; e += 64) {
        const float dv = img[box_voxel(e, b.w, b.h, b.x, b.y, b.z, w, plane)] - m;
        out[e] = dv * dv;
    }
}

// OcTree_.applyGATtree (:471-485): 0 where v < -ref/g0 (thr), else
// (2/g0) sqrt(g0 v + ref), in double, stored as float.
__global__ __launch_bounds__(256) void k_gat_tree3d(float *__restrict__ px, int64_t n, double g, double ref,
                                                    double thr, double two_over_g)
{
#pragma clang fp contract(off)
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        double v = (double)px[i];
        if (v < thr) v = 0.0;
        else v = two_over_g * sqrt(g * v + ref);
        px[i] = (float)v;
    }
}

}  // namespace

int octree_gat_dev(sred_ctx *ctx, const float *d_vol, int w, int h, int d, int min_size, float alpha, int max_iter,
                   float frac, double *g0_edc, int *n_leaves, int cap, int32_t *boxes, float *means, float *vars)
{
    if (w < 1 || h < 1 || d < 1 || (int64_t)w * h * d >= (1LL << 31) || h > 65535 || d > 65535 || min_size < 1 ||
        !(alpha >= 0.0f && alpha <= 1.0f) || !(frac >= 0.0f && frac <= 1.0f) || cap < 0) {
        set_error("octree GAT: needs w, h, d >= 1 (w*h*d < 2^31, h and d <= 65535), min_size >= 1, alpha and the LTS "
                  "fraction in [0, 1]");
        return SRED_ERR_ARG;
    }
    // level geometry: nodes of level k are W[k] x H[k] x D[k]; level k+1 exists while k's can divide
    OcLevels L{};
    int W = w, H = h, D = d, depth = 0;
    std::vector<int> LW{W}, LH{H}, LD{D};
    while (W > min_size && H > min_size && D > min_size) {
        W /= 2;
        H /= 2;
        D /= 2;
        LW.push_back(W);
        LH.push_back(H);
        LD.push_back(D);
        depth++;
    }
    L.nlev = std::max(depth, 1);  // the root's criterion is evaluated even when it cannot divide
    if (L.nlev > kOcMaxLevels - 1) {
        set_error("octree GAT: too many levels");
        return SRED_ERR_ARG;
    }
    if ((int64_t)w * h <= 1) {
        set_error("octree GAT: a one-pixel slice has no F test (FDistribution(0, 0))");
        return SRED_ERR_ARG;
    }
    long long nodes = 0;
    for (int k = 0; k < L.nlev; k++) {
        L.W[k] = LW[k];
        L.H[k] = LH[k];
        L.D[k] = LD[k];
        L.first[k] = nodes;
        nodes += 1LL << (3 * k);
    }
    L.first[L.nlev] = nodes;
    if (nodes > (1LL << 27)) {
        set_error("octree GAT: too many tree nodes (raise min_size)");
        return SRED_ERR_ARG;
    }
    const int64_t nvox = (int64_t)w * h * d;
    SRED_TRY(ctx->qt_res.ensure(sizeof(float) * nvox + 256));
    SRED_TRY(ctx->qt_nodes.ensure(sizeof(float2) * nodes + 256));
    const float scale = (float)(1.0 / std::sqrt(56.0));  // (float)(1/sqrt(l*l + l)), l = 7 (OcTree_.java:241)
    hipLaunchKernelGGL(k_oc_residuals, dim3((w + 63) / 64, (h + 3) / 4, d), dim3(256), 0, ctx->stream, d_vol, w, h, d,
                       scale, ctx->qt_res.as<float>());
    SRED_TRY(time_begin(ctx));
    hipLaunchKernelGGL(k_oc_criteria, dim3((unsigned)(2 * nodes)), dim3(64), 0, ctx->stream, d_vol, ctx->qt_res.as<float>(),
                       w, h, L, ctx->qt_nodes.as<float2>());
    SRED_TRY(time_end(ctx, nvox * L.nlev));
    SRED_HIP_CHECK(hipGetLastError());
    std::vector<float2> st((size_t)nodes);
    SRED_HIP_CHECK(hipMemcpyAsync(st.data(), ctx->qt_nodes.ptr, sizeof(float2) * nodes, hipMemcpyDeviceToHost,
                                  ctx->stream));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    // Fisher thresholds, one per level, with (width*height - 1) degrees of freedom (:282-296, :225)
    std::vector<float> thr(L.nlev);
    for (int k = 0; k < L.nlev; k++) {
        const double dof = (double)L.W[k] * L.H[k] - 1.0;
        thr[k] = (float)f_quantile(dof, dof, (double)alpha);
    }
    // pre-order walk, children x fastest then y then z (splitNode, :197-213)
    std::vector<Box> leaves;
    struct Item {
        int k, i, j, l, x, y, z;
    };
    std::vector<Item> stack{{0, 0, 0, 0, 0, 0, 0}};
    while (!stack.empty()) {
        const Item it = stack.back();
        stack.pop_back();
        if ((int64_t)LW[it.k] * LH[it.k] <= 1) {  // its criterion would build FDistribution(0, 0)
            set_error("octree GAT: a visited node has one pixel per slice (FDistribution(0, 0))");
            return SRED_ERR_ARG;
        }
        bool divide = false;
        if (it.k < L.nlev && it.k < depth) {
            const long long s = 1LL << it.k;
            const float2 v = st[(size_t)(L.first[it.k] + ((long long)it.l * s + it.j) * s + it.i)];
            const float ratio = java_minf(v.x, v.y) / java_maxf(v.x, v.y);
            divide = ratio < thr[it.k];
        }
        if (!divide) {
            leaves.push_back(Box{it.x, it.y, it.z, LW[it.k], LH[it.k], LD[it.k]});
            continue;
        }
        const int k1 = it.k + 1, hw = LW[k1], hh = LH[k1], hd = LD[k1];
        for (int c = 7; c >= 0; c--) {  // pushed in reverse: popped 0, 1, ..., 7
            const int bx = c & 1, by = (c >> 1) & 1, bz = c >> 2;
            stack.push_back({k1, 2 * it.i + bx, 2 * it.j + by, 2 * it.l + bz, it.x + bx * hw, it.y + by * hh,
                             it.z + bz * hd});
        }
    }
    const int nl = (int)leaves.size();
    if (n_leaves) *n_leaves = nl;
    std::vector<int> off(nl + 1, 0);
    for (int i = 0; i < nl; i++) off[i + 1] = off[i] + leaves[i].w * leaves[i].h * leaves[i].d;
    const int nseg = off[nl];

    SRED_TRY(ctx->qt_leaves.ensure(sizeof(Box) * nl + sizeof(int) * (nl + 1) + 2 * sizeof(float) * nl + 1024));
    Box *d_leaves = ctx->qt_leaves.as<Box>();
    int *d_off = reinterpret_cast<int *>(d_leaves + nl);
    float *d_means = reinterpret_cast<float *>(d_off + nl + 1), *d_vars = d_means + nl;
    SRED_HIP_CHECK(hipMemcpyAsync(d_leaves, leaves.data(), sizeof(Box) * nl, hipMemcpyHostToDevice, ctx->stream));
    SRED_HIP_CHECK(hipMemcpyAsync(d_off, off.data(), sizeof(int) * (nl + 1), hipMemcpyHostToDevice, ctx->stream));
    hipLaunchKernelGGL(k_oc_robust_mean, dim3(nl), dim3(64), 0, ctx->stream, d_vol, w, h, d_leaves, max_iter, d_means);
    SRED_TRY(ctx->qt_seg.ensure(2 * sizeof(float) * nseg + 256));
    float *seg_in = ctx->qt_seg.as<float>(), *seg_out = seg_in + nseg;
    hipLaunchKernelGGL(k_oc_lts_residuals, dim3(nl), dim3(64), 0, ctx->stream, d_vol, w, h, d_leaves, d_off, d_means,
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
        if (boxes) {
            int32_t *r = boxes + 6 * (size_t)i;
            r[0] = leaves[i].x, r[1] = leaves[i].y, r[2] = leaves[i].z;
            r[3] = leaves[i].w, r[4] = leaves[i].h, r[5] = leaves[i].d;
        }
        if (means) means[i] = mv[i];
        if (vars) vars[i] = mv[nl + i];
    }
    if (nl < 3) {
        set_error("octree GAT: fewer than 3 leaves for the regression (SimpleRegression.regress needs 3 points)");
        return SRED_ERR_ARG;
    }
    std::vector<double> x(nl), y(nl);
    for (int i = 0; i < nl; i++) {
        x[i] = (double)mv[i];
        y[i] = (double)mv[nl + i];
    }
    double s, c;
    simple_regression_host(x.data(), y.data(), nl, &s, &c);
    g0_edc[0] = (double)(float)s;  // performLinearRegression returns floats (:465-466)
    g0_edc[1] = (double)(float)c;
    return SRED_OK;
}

int gat_tree3d_dev(sred_ctx *ctx, float *d_vol, int64_t n, float g0, float edc)
{
#pragma clang fp contract(off)
    if (n <= 0) return SRED_OK;
    const double g = (double)g0;
    const double ref = 3.0 / 8.0 * g * g + (double)edc;
    const int blocks = (int)std::min<int64_t>((n + 255) / 256, 8192);
    hipLaunchKernelGGL(k_gat_tree3d, dim3(blocks), dim3(256), 0, ctx->stream, d_vol, n, g, ref, -ref / g, 2.0 / g);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

}  // namespace sred
