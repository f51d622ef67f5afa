// stats.hip -- per-block statistics, noise variance, relevance mask and the
// small map operations of the host finalize (normalise, mask, invert).
//
// Float semantics: these kernels reproduce Java / OpenCL float32 results
// bit-for-bit, so every body that accumulates runs with FP contraction off
// (no fused multiply-add) and uses IEEE division and square root (HIP's
// default correctly rounded f32 div/sqrt).
#include <float.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "sred_internal.h"

namespace sred {

// ---------------------------------------------------------------------------
// Local statistics: kernelGetLocalStatistics2D.cl:9-55.
// 16x16 pixels per workgroup; the (16+2brh) x (16+2brw) input tile is staged
// once in LDS with coalesced loads, then every pixel sums its elliptical
// block from LDS in the reference's row-major order (sequential float sums,
// so means/stds equal the reference's float results exactly).
// ---------------------------------------------------------------------------
constexpr int kStatTile = 16;

// kVar: stds receives the variance itself (kernelGetRelevanceMask2D.cl), not its root.
template <bool kVar>
__global__ __launch_bounds__(256) void k_local_stats(const float *__restrict__ img, int w, int h, int brw, int brh,
                                                     const int *__restrict__ offs, int K, float *__restrict__ means,
                                                     float *__restrict__ stds)
{
#pragma clang fp contract(off)
    extern __shared__ float tile[];
    const int tw = kStatTile + 2 * brw, th = kStatTile + 2 * brh;
    const int x0 = blockIdx.x * kStatTile, y0 = blockIdx.y * kStatTile;
    for (int i = threadIdx.x; i < tw * th; i += blockDim.x) {
        int ly = i / tw, lx = i - ly * tw;
        int gy = y0 - brh + ly, gx = x0 - brw + lx;
        tile[i] = (gx >= 0 && gx < w && gy >= 0 && gy < h) ? img[(int64_t)gy * w + gx] : 0.0f;
    }
    __syncthreads();
    const int tx = threadIdx.x % kStatTile, ty = threadIdx.x / kStatTile;
    const int gx = x0 + tx, gy = y0 + ty;
    if (gx >= w || gy >= h) return;
    const int64_t o = (int64_t)gy * w + gx;
    if (gx < brw || gx >= w - brw || gy < brh || gy >= h - brh) {
        means[o] = 0.0f;
        stds[o] = 0.0f;
        return;
    }
    // offs[k] = dy*tw + dx relative to the block centre in the LDS tile
    const float *c = tile + (ty + brh) * tw + (tx + brw);
    float m = 0.0f;
    for (int k = 0; k < K; k++) m += c[offs[k]];
    m /= (float)K;
    float v = 0.0f;
    for (int k = 0; k < K; k++) {
        float d = c[offs[k]] - m;
        v += d * d;
    }
    v /= (float)(K - 1);
    means[o] = m;
    stds[o] = kVar ? v : sqrtf(v);
}

int launch_local_stats(sred_ctx *ctx, const Geometry &g, const float *d_img, float *d_means, float *d_stds,
                       bool var_out)
{
    const int tw = kStatTile + 2 * g.brw;
    SRED_TRY(upload_offsets(ctx, g, tw, &ctx->offs, ctx->offs_key));
    dim3 grid((g.w + kStatTile - 1) / kStatTile, (g.h + kStatTile - 1) / kStatTile);
    size_t lds = sizeof(float) * (size_t)(kStatTile + 2 * g.brw) * (kStatTile + 2 * g.brh);
    if (var_out)
        hipLaunchKernelGGL(k_local_stats<true>, grid, dim3(256), lds, ctx->stream, d_img, g.w, g.h, g.brw, g.brh,
                           ctx->offs.as<int>(), g.K, d_means, d_stds);
    else
        hipLaunchKernelGGL(k_local_stats<false>, grid, dim3(256), lds, ctx->stream, d_img, g.w, g.h, g.brw, g.brh,
                           ctx->offs.as<int>(), g.K, d_means, d_stds);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

// ---------------------------------------------------------------------------
// Noise variance: Utils.getMeanNoiseVar2D (Utils.java:790-837) with
// getMeanAndVarBlock2D (Utils.java:854-900).  One thread per tile runs the
// reference's one-pass float loop; the lowest-3% mean is taken on the host
// exactly as the Java code does (sort, sequential float sum).  With T =
// double, the tile sums are the double ones of RelevanceMask2D_'s
// getMeanAndVarBlock (RelevanceMask2D_.java:401-419), the variance rounded
// to float.
// ---------------------------------------------------------------------------
template <class T>
__global__ void k_tile_var(const float *__restrict__ img, int w, int tbw, int tbh, int nbx, int nb,
                           float *__restrict__ out, int *__restrict__ bad)
{
#pragma clang fp contract(off)
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= nb) return;
    int bx = t % nbx, by = t / nbx;
    T m = 0, sq = 0;
    int nonfinite = 0;
    for (int y = by * tbh; y < (by + 1) * tbh; y++)
        for (int x = bx * tbw; x < (bx + 1) * tbw; x++) {
            float v = img[(int64_t)y * w + x];
            nonfinite |= !isfinite(v);
            m += v;
            sq += v * v;  // the float square in both forms
        }
    int n = tbw * tbh;
    m = m / (T)n;
    const float var = (float)(sq / (T)n - m * m);
    if (nonfinite || isnan(var)) atomicOr(bad, 1);
    out[t] = var;
}

// Tile count of Utils.getMeanNoiseVar2D: 8x8 tiles, 16x16 above the CIF
// size (Utils.java:793-801).
int noise_tiles(int w, int h, int *tbw, int *tbh)
{
    *tbw = 8;
    *tbh = 8;
    if ((int64_t)w * h > 352 * 288) { *tbw = 16; *tbh = 16; }
    const int nb = (w / *tbw) * (h / *tbh);
    if (nb <= 0) {
        set_error("image smaller than one noise tile (the reference indexes an empty variance array)");
        return -1;
    }
    return nb;
}

int noise_var_enqueue(sred_ctx *ctx, DevBuf &tiles, const float *d_img, int w, int h, float *h_tiles, bool double_sums)
{
    int tbw, tbh;
    const int nb = noise_tiles(w, h, &tbw, &tbh);
    if (nb < 0) return SRED_ERR_ARG;
    const int nbx = w / tbw;
    SRED_TRY(tiles.ensure(sizeof(float) * nb + 64));
    float *d_var = tiles.as<float>();
    int *d_bad = reinterpret_cast<int *>(d_var + nb);
    SRED_TRY(zero_dev(ctx, d_bad, sizeof(int)));
    if (double_sums)
        hipLaunchKernelGGL(k_tile_var<double>, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, d_img, w, tbw, tbh,
                           nbx, nb, d_var, d_bad);
    else
        hipLaunchKernelGGL(k_tile_var<float>, dim3((nb + 255) / 256), dim3(256), 0, ctx->stream, d_img, w, tbw, tbh,
                           nbx, nb, d_var, d_bad);
    SRED_HIP_CHECK(hipGetLastError());
    SRED_HIP_CHECK(hipMemcpyAsync(h_tiles, d_var, sizeof(float) * (nb + 1), hipMemcpyDeviceToHost, ctx->stream));
    return SRED_OK;
}

// The host part of Utils.getMeanNoiseVar2D (Utils.java:818-836) on the tile
// variances and the non-finite flag (h_tiles[nb]) the enqueue copied back;
// sorts h_tiles in place.
int noise_var_finish(float *h_tiles, int nb, float *noise_var)
{
#pragma clang fp contract(off)
    int bad;
    memcpy(&bad, &h_tiles[nb], sizeof(int));
    if (bad) {
        set_error("Input contains invalid (NaN or Infinite) values.");
        return SRED_ERR_NONFINITE;
    }
    std::sort(h_tiles, h_tiles + nb);
    int nvars = (int)(0.03f * (float)nb + 1.0f);
    float nv = 0.0f;
    for (int i = 0; i < nvars && i < nb; i++) nv += h_tiles[i];
    nv = fabsf(nv / (float)nvars);
    nv = (1.0f + 0.001f * (nv - 40.0f)) * nv;
    *noise_var = nv;
    return SRED_OK;
}

int noise_var_dev(sred_ctx *ctx, const float *d_img, int w, int h, float *noise_var, bool double_sums)
{
    int tbw, tbh;
    const int nb = noise_tiles(w, h, &tbw, &tbh);
    if (nb < 0) return SRED_ERR_ARG;
    ctx->host_tile_var.resize(nb + 1);
    SRED_TRY(noise_var_enqueue(ctx, ctx->tile_var, d_img, w, h, ctx->host_tile_var.data(), double_sums));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return noise_var_finish(ctx->host_tile_var.data(), nb, noise_var);
}

// ---------------------------------------------------------------------------
// Relevance mask (Utils.java:944-968): 0 if std*std <= thr else 1 on the
// interior, 0 on the border; nPixels counts ones (GlobalRepetition2D_.java:
// 133-140).  Java accumulates nPixels as float += 1.0f, which saturates at
// 2^24; the count is clamped the same way.
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(256) void k_relevance_mask(const float *__restrict__ stds, int w, int h, int brw,
                                                        int brh, float thr, float *__restrict__ mask,
                                                        unsigned long long *__restrict__ count)
{
#pragma clang fp contract(off)
    int64_t n = (int64_t)w * h;
    int local = 0;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        int x = (int)(i % w), y = (int)(i / w);
        float m = 0.0f;
        if (x >= brw && x < w - brw && y >= brh && y < h - brh) {
            float s = stds[i];
            m = (s * s <= thr) ? 0.0f : 1.0f;
        }
        mask[i] = m;
        local += (m > 0.0f);
    }
    __shared__ int red[256];
    red[threadIdx.x] = local;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) red[threadIdx.x] += red[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0 && red[0]) atomicAdd(count, (unsigned long long)red[0]);
}

int relevance_mask_enqueue(sred_ctx *ctx, const Geometry &g, const float *d_stds, float threshold, float *d_mask,
                           unsigned long long *d_count, unsigned long long *h_count)
{
    SRED_TRY(zero_dev(ctx, d_count, sizeof(unsigned long long)));
    int blocks = (int)std::min<int64_t>((g.n() + 255) / 256, 4096);
    hipLaunchKernelGGL(k_relevance_mask, dim3(blocks), dim3(256), 0, ctx->stream, d_stds, g.w, g.h, g.brw, g.brh,
                       threshold, d_mask, d_count);
    SRED_HIP_CHECK(hipGetLastError());
    SRED_HIP_CHECK(hipMemcpyAsync(h_count, d_count, sizeof *h_count, hipMemcpyDeviceToHost, ctx->stream));
    return SRED_OK;
}

// nPixels as Java counts it: float += 1.0f, which stops growing at 2^24.
float n_pixels_of(unsigned long long count) { return (float)std::min<unsigned long long>(count, 16777216ull); }

int relevance_mask_dev(sred_ctx *ctx, const Geometry &g, const float *d_stds, float threshold, float *d_mask,
                       float *n_pixels)
{
    SRED_TRY(ctx->misc.ensure(kMiscBytes));
    unsigned long long cnt = 0;
    SRED_TRY(relevance_mask_enqueue(ctx, g, d_stds, threshold, d_mask, ctx->misc.as<unsigned long long>(), &cnt));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *n_pixels = n_pixels_of(cnt);
    return SRED_OK;
}

// ---------------------------------------------------------------------------
// Map operations of the host finalize.
// ---------------------------------------------------------------------------
__device__ __forceinline__ bool interior(int64_t i, int w, int h, int brw, int brh)
{
    const int ii = (int)i;  // images hold fewer than 2^31 pixels
    const int y = ii / w, x = ii - y * w;
    return x >= brw && x < w - brw && y >= brh && y < h - brh;
}

__global__ void k_minmax_init(unsigned int *mm, float init_max)
{
    mm[0] = f2ord(FLT_MAX);
    mm[1] = f2ord(init_max);
    mm[2] = 0;  // non-binary flag
}

// Java Math.min/Math.max over the interior (mask > 0 where a mask is given);
// order-independent, so a parallel reduction gives the sequential result.
// Two stages without atomics: each workgroup writes its (min, max, non-binary)
// to part[3 * blockIdx.x ..], then k_minmax_final folds them into mm (which
// holds the initial values).
__global__ __launch_bounds__(256) void k_minmax(const float *__restrict__ map, const float *__restrict__ mask, int w,
                                                int h, int brw, int brh, int d, int brd,
                                                unsigned int *__restrict__ part)
{
    unsigned int lo = 0xffffffffu, hi = 0u;
    int nonbinary = 0;
    // whole interior rows per workgroup (no per-pixel index division); a
    // volume's rows run slice by slice
    const int ih = h - 2 * brh, rows = ih * (d - 2 * brd);
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const int z = brd + r / ih, y = brh + r % ih;
        const int64_t row = ((int64_t)z * h + y) * w;
        for (int x = brw + threadIdx.x; x < w - brw; x += blockDim.x) {
            const int64_t i = row + x;
            if (mask) {
                float m = mask[i];
                if (m != 0.0f && m != 1.0f) nonbinary = 1;
                if (!(m > 0.0f)) continue;
            }
            unsigned int o = f2ord(map[i]);
            lo = min(lo, o);
            hi = max(hi, o);
        }
    }
    __shared__ unsigned int slo[256], shi[256];
    __shared__ int snb;
    if (threadIdx.x == 0) snb = 0;
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    __syncthreads();
    if (nonbinary) atomicOr(&snb, 1);
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            slo[threadIdx.x] = min(slo[threadIdx.x], slo[threadIdx.x + s]);
            shi[threadIdx.x] = max(shi[threadIdx.x], shi[threadIdx.x + s]);
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        part[3 * blockIdx.x] = slo[0];
        part[3 * blockIdx.x + 1] = shi[0];
        part[3 * blockIdx.x + 2] = snb ? 1u : 0u;
    }
}

__device__ __forceinline__ void k_minmax_final_body(const unsigned int *__restrict__ part, int nparts,
                                                    unsigned int *__restrict__ mm);
__global__ __launch_bounds__(256) void k_minmax_final(const unsigned int *__restrict__ part, int nparts,
                                                      unsigned int *__restrict__ mm)
{
    k_minmax_final_body(part, nparts, mm);
}
__device__ __forceinline__ void k_minmax_final_body(const unsigned int *__restrict__ part, int nparts,
                                                    unsigned int *__restrict__ mm)
{
    unsigned int lo = 0xffffffffu, hi = 0u, nb = 0u;
    for (int i = threadIdx.x; i < nparts; i += 256) {
        lo = min(lo, part[3 * i]);
        hi = max(hi, part[3 * i + 1]);
        nb |= part[3 * i + 2];
    }
    __shared__ unsigned int slo[256], shi[256], snb[256];
    slo[threadIdx.x] = lo;
    shi[threadIdx.x] = hi;
    snb[threadIdx.x] = nb;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if (threadIdx.x < s) {
            slo[threadIdx.x] = min(slo[threadIdx.x], slo[threadIdx.x + s]);
            shi[threadIdx.x] = max(shi[threadIdx.x], shi[threadIdx.x + s]);
            snb[threadIdx.x] |= snb[threadIdx.x + s];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        mm[0] = min(mm[0], slo[0]);
        mm[1] = max(mm[1], shi[0]);
        mm[2] |= snb[0];
    }
}

// mm = (min, max, non-binary) of the initial values folded with nparts
// partials (one launch: initialisation and final reduction together).
__global__ __launch_bounds__(256) void k_minmax_from_parts(const unsigned int *__restrict__ part, int nparts,
                                                           float init_max, unsigned int *__restrict__ mm)
{
    if (threadIdx.x == 0) {
        mm[0] = f2ord(FLT_MAX);
        mm[1] = f2ord(init_max);
        mm[2] = 0;
    }
    __syncthreads();
    k_minmax_final_body(part, nparts, mm);
}

// Utils.java:1014-1028: (x - min) / (max - min + EPSILON) on the selected pixels.
__global__ __launch_bounds__(256) void k_remap(float *__restrict__ map, const float *__restrict__ mask, int w, int h,
                                               int brw, int brh, int d, int brd, const unsigned int *__restrict__ mm)
{
#pragma clang fp contract(off)
    if (mm[2]) return;  // non-binary mask: the reference throws before remapping
    const float mn = ord2f(mm[0]), mx = ord2f(mm[1]);
    const float den = mx - mn + kEps;
    const int ih = h - 2 * brh, rows = ih * (d - 2 * brd);
    for (int r = blockIdx.x; r < rows; r += gridDim.x) {
        const int z = brd + r / ih, y = brh + r % ih;
        const int64_t row = ((int64_t)z * h + y) * w;
        for (int x = brw + threadIdx.x; x < w - brw; x += blockDim.x) {
            const int64_t i = row + x;
            if (mask && !(mask[i] > 0.0f)) continue;
            map[i] = (map[i] - mn) / den;
        }
    }
}

static int grid_for(const Geometry &g) { return (int)std::min<int64_t>((g.n() + 255) / 256, 8192); }
constexpr int kMinmaxParts = 1024;  // workgroups of k_minmax (grid-stride); kMiscBytes holds their partials
constexpr size_t kMiscPartOffset = 4096;
// one workgroup per interior row, at most kMinmaxParts (partial slots)
static int grid_minmax(const Geometry &g)
{
    return std::max(1, std::min((g.h - 2 * g.brh) * (g.d - 2 * g.brd), kMinmaxParts));
}

// min/max (and the non-binary flag) of map over the selected pixels into mm[0..2]
static int minmax_dev(sred_ctx *ctx, const Geometry &g, const float *d_map, const float *d_mask, unsigned int *mm)
{
    unsigned int *part = reinterpret_cast<unsigned int *>(ctx->misc.as<char>() + kMiscPartOffset);
    const int grid = grid_minmax(g);
    hipLaunchKernelGGL(k_minmax, dim3(grid), dim3(256), 0, ctx->stream, d_map, d_mask, g.w, g.h, g.brw, g.brh, g.d,
                       g.brd, part);
    hipLaunchKernelGGL(k_minmax_final, dim3(1), dim3(256), 0, ctx->stream, part, grid, mm);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

int normalize_map_dev(sred_ctx *ctx, const Geometry &g, float *d_map, const float *d_mask, int init_max_min_value)
{
    SRED_TRY(ctx->misc.ensure(kMiscBytes));
    unsigned int *mm = reinterpret_cast<unsigned int *>(ctx->misc.as<char>() + 1024);
    float init_max = init_max_min_value ? 1.4e-45f : -FLT_MAX;  // Float.MIN_VALUE (CLUtils.java:940)
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(1), 0, ctx->stream, mm, init_max);
    SRED_TRY(minmax_dev(ctx, g, d_map, d_mask, mm));
    hipLaunchKernelGGL(k_remap, dim3(std::max(1, std::min((g.h - 2 * g.brh) * (g.d - 2 * g.brd), 8192))), dim3(256), 0,
                       ctx->stream, d_map, d_mask, g.w, g.h, g.brw, g.brh, g.d, g.brd, mm);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

int normalize_map_from_parts_dev(sred_ctx *ctx, const Geometry &g, float *d_map, const unsigned int *d_parts,
                                 int nparts, int init_max_min_value)
{
    SRED_TRY(ctx->misc.ensure(kMiscBytes));
    unsigned int *mm = reinterpret_cast<unsigned int *>(ctx->misc.as<char>() + 1024);
    const float init_max = init_max_min_value ? 1.4e-45f : -FLT_MAX;  // Float.MIN_VALUE (CLUtils.java:940)
    hipLaunchKernelGGL(k_minmax_from_parts, dim3(1), dim3(256), 0, ctx->stream, d_parts, nparts, init_max, mm);
    hipLaunchKernelGGL(k_remap, dim3(std::max(1, std::min((g.h - 2 * g.brh) * (g.d - 2 * g.brd), 8192))), dim3(256), 0,
                       ctx->stream, d_map, nullptr, g.w, g.h, g.brw, g.brh, g.d, g.brd, mm);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

// The same for nframes maps of a stack (frame z at d_maps + z*w*h, partials
// at d_parts + 3*nparts*z): one workgroup per frame, then one remap launch.
__global__ __launch_bounds__(256) void k_frames_minmax_from_parts(const unsigned int *__restrict__ part, int nparts,
                                                                  float init_max, unsigned int *__restrict__ mm)
{
    unsigned int *m = mm + 4 * blockIdx.x;
    if (threadIdx.x == 0) {
        m[0] = f2ord(FLT_MAX);
        m[1] = f2ord(init_max);
        m[2] = 0;
    }
    __syncthreads();
    k_minmax_final_body(part + 3 * (size_t)nparts * blockIdx.x, nparts, m);
}

__global__ __launch_bounds__(256) void k_frames_remap(float *__restrict__ maps, int w, int h, int brw, int brh,
                                                      const unsigned int *__restrict__ mm)
{
#pragma clang fp contract(off)
    const unsigned int *m = mm + 4 * blockIdx.y;
    const float mn = ord2f(m[0]), mx = ord2f(m[1]);
    const float den = mx - mn + kEps;
    float *map = maps + (int64_t)blockIdx.y * w * h;
    for (int y = brh + blockIdx.x; y < h - brh; y += gridDim.x) {
        const int64_t row = (int64_t)y * w;
        for (int x = brw + threadIdx.x; x < w - brw; x += blockDim.x) map[row + x] = (map[row + x] - mn) / den;
    }
}

int normalize_frames_from_parts_dev(sred_ctx *ctx, const Geometry &g, float *d_maps, int nframes,
                                    const unsigned int *d_parts, int nparts, int init_max_min_value)
{
    if (nframes < 1 || nframes > 65535) {
        set_error("invalid frame count");
        return SRED_ERR_ARG;
    }
    SRED_TRY(ctx->seg_mm.ensure(sizeof(unsigned int) * 4 * (size_t)nframes + 256));
    unsigned int *mm = ctx->seg_mm.as<unsigned int>();
    const float init_max = init_max_min_value ? 1.4e-45f : -FLT_MAX;  // Float.MIN_VALUE (CLUtils.java:940)
    hipLaunchKernelGGL(k_frames_minmax_from_parts, dim3(nframes), dim3(256), 0, ctx->stream, d_parts, nparts,
                       init_max, mm);
    const int rows = std::max(1, g.h - 2 * g.brh);
    const int gx = std::max(1, std::min(rows, std::max(1, 8192 / nframes)));
    hipLaunchKernelGGL(k_frames_remap, dim3(gx, nframes), dim3(256), 0, ctx->stream, d_maps, g.w, g.h, g.brw, g.brh,
                       mm);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

int mask_is_binary_dev(sred_ctx *ctx, const Geometry &g, const float *d_mask, int *is_binary)
{
    SRED_TRY(ctx->misc.ensure(kMiscBytes));
    unsigned int *mm = reinterpret_cast<unsigned int *>(ctx->misc.as<char>() + 2048);
    hipLaunchKernelGGL(k_minmax_init, dim3(1), dim3(1), 0, ctx->stream, mm, -FLT_MAX);
    SRED_TRY(minmax_dev(ctx, g, d_mask, d_mask, mm));
    SRED_HIP_CHECK(hipGetLastError());
    unsigned int flag = 0;
    SRED_HIP_CHECK(hipMemcpyAsync(&flag, mm + 2, sizeof flag, hipMemcpyDeviceToHost, ctx->stream));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    *is_binary = flag == 0;
    return SRED_OK;
}

__global__ void k_apply_mask(float *__restrict__ map, const float *__restrict__ mask, int64_t n)
{
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        map[i] *= mask[i];
}

int apply_mask_dev(sred_ctx *ctx, const Geometry &g, float *d_map, const float *d_mask)
{
    hipLaunchKernelGGL(k_apply_mask, dim3(grid_for(g)), dim3(256), 0, ctx->stream, d_map, d_mask, g.n());
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

__global__ void k_interior_ones(float *__restrict__ mask, int w, int h, int brw, int brh)
{
    int64_t n = (int64_t)w * h;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        mask[i] = interior(i, w, h, brw, brh) ? 1.0f : 0.0f;
}

int interior_ones_mask_dev(sred_ctx *ctx, const Geometry &g, float *d_mask)
{
    hipLaunchKernelGGL(k_interior_ones, dim3(grid_for(g)), dim3(256), 0, ctx->stream, d_mask, g.w, g.h, g.brw, g.brh);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

__global__ void k_invert_masked(float *__restrict__ map, const float *__restrict__ mask, int w, int h, int brw, int brh)
{
    int64_t n = (int64_t)w * h;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x) {
        if (!interior(i, w, h, brw, brh) || !(mask[i] > 0.0f)) continue;
        float v = map[i];
        map[i] = (v == 0.0f) ? 1.0f : 1.0f / v;
    }
}

int invert_masked_dev(sred_ctx *ctx, const Geometry &g, float *d_map, const float *d_mask)
{
    hipLaunchKernelGGL(k_invert_masked, dim3(grid_for(g)), dim3(256), 0, ctx->stream, d_map, d_mask, g.w, g.h, g.brw,
                       g.brh);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

__global__ void k_one_minus(float *__restrict__ map, int w, int h, int brw, int brh)
{
    int64_t n = (int64_t)w * h;
    for (int64_t i = blockIdx.x * (int64_t)blockDim.x + threadIdx.x; i < n; i += (int64_t)gridDim.x * blockDim.x)
        if (interior(i, w, h, brw, brh)) map[i] = 1.0f - map[i];
}

int one_minus_dev(sred_ctx *ctx, const Geometry &g, float *d_map)
{
    hipLaunchKernelGGL(k_one_minus, dim3(grid_for(g)), dim3(256), 0, ctx->stream, d_map, g.w, g.h, g.brw, g.brh);
    SRED_HIP_CHECK(hipGetLastError());
    return SRED_OK;
}

}  // namespace sred
