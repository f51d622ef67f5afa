// api.hip -- the extern "C" boundary of libsred_hip.so (declared in
// include/sred.h) and the host-side helpers it needs.
//
// The host-pointer entry points reproduce the CLUtils methods a JNI binding
// replaces; the _dev entry points run the same pipelines on device buffers.
// Host-side mirrors of Utils (noise variance, relevance mask, normalisation,
// reference block) are plain C++ with Java float semantics, compiled without
// FP contraction (the file is built with -ffp-contract=off for the host).
#include <float.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "sred_internal.h"

#define SRED_VERSION 100

namespace sred {

static thread_local std::string g_last_error;

void set_error(const std::string &msg) { g_last_error = msg; }

int DevBuf::ensure(size_t n)
{
    if (n <= bytes && ptr) return SRED_OK;
    if (fixed) {
        set_error("lent scratch region too small (" + std::to_string(bytes) + " < " + std::to_string(n) + " bytes)");
        return SRED_ERR_ARG;
    }
    if (ptr) {
        hipError_t e = hipFree(ptr);
        ptr = nullptr;
        bytes = 0;
        if (e != hipSuccess) {
            set_error(std::string("hipFree: ") + hipGetErrorString(e));
            return SRED_ERR_HIP;
        }
    }
    size_t want = std::max<size_t>(n, 256);
    hipError_t e = hipMalloc(&ptr, want);
    if (e != hipSuccess) {
        ptr = nullptr;
        set_error(std::string("hipMalloc(") + std::to_string(want) + "): " + hipGetErrorString(e));
        return SRED_ERR_HIP;
    }
    bytes = want;
    return SRED_OK;
}

void DevBuf::release()
{
    if (ptr && !fixed) (void)hipFree(ptr);
    fixed = false;
    ptr = nullptr;
    bytes = 0;
}

// GlobalRepetition2D_.java:102-111 / CLUtils.java:391-400 / kernel predicate.
int footprint(int bw, int bh, std::vector<int> *dx, std::vector<int> *dy)
{
#pragma clang fp contract(off)
    const int brw = bw / 2, brh = bh / 2;
    int k = 0;
    if (dx) { dx->clear(); dy->clear(); }
    for (int y = 0; y < bh; y++)
        for (int x = 0; x < bw; x++) {
            float fx = (float)(x - brw), fy = (float)(y - brh);
            if (((fx * fx) / (float)(brw * brw)) + ((fy * fy) / (float)(brh * brh)) <= 1.0f) {
                if (dx) { dx->push_back(x - brw); dy->push_back(y - brh); }
                k++;
            }
        }
    return k;
}

// Ellipsoid footprint of the 3D kernels (kernelGetLocalStatistics3D.cl,
// kernelGetGlobalPearson3D.cl; GlobalRepetition3D_.java:101-112): z, then y,
// then x, with the float predicate dx^2/brw^2 + dy^2/brh^2 + dz^2/brd^2 <= 1
// summed left to right.
int footprint3d(int bw, int bh, int bd, std::vector<int> *dx, std::vector<int> *dy, std::vector<int> *dz)
{
#pragma clang fp contract(off)
    const int brw = bw / 2, brh = bh / 2, brd = bd / 2;
    int k = 0;
    if (dx) { dx->clear(); dy->clear(); dz->clear(); }
    for (int z = 0; z < bd; z++)
        for (int y = 0; y < bh; y++)
            for (int x = 0; x < bw; x++) {
                const float fx = (float)(x - brw), fy = (float)(y - brh), fz = (float)(z - brd);
                if (((fx * fx) / (float)(brw * brw)) + ((fy * fy) / (float)(brh * brh)) +
                        ((fz * fz) / (float)(brd * brd)) <=
                    1.0f) {
                    if (dx) { dx->push_back(x - brw); dy->push_back(y - brh); dz->push_back(z - brd); }
                    k++;
                }
            }
    return k;
}

int make_geometry3d(int w, int h, int d, int bw, int bh, int bd, Geometry *g)
{
    if (w <= 0 || h <= 0 || d <= 0) {
        set_error("volume dimensions must be positive");
        return SRED_ERR_ARG;
    }
    if (bw < 3 || bh < 3 || bd < 3 || bw % 2 == 0 || bh % 2 == 0 || bd % 2 == 0) {
        set_error("Block dimensions must be odd, and the block depth at least 3. Please try again.");
        return SRED_ERR_ARG;
    }
    if (bw > SRED_MAX_BLOCK_SIDE || bh > SRED_MAX_BLOCK_SIDE || bd > SRED_MAX_BLOCK_SIDE) {
        set_error("block side larger than SRED_MAX_BLOCK_SIDE");
        return SRED_ERR_ARG;
    }
    if ((int64_t)w * h * d >= (int64_t)1 << 31 || h > 65535 || d > 65535) {
        set_error("volume larger than 2^31 voxels, or more than 65535 rows or slices");
        return SRED_ERR_ARG;
    }
    g->w = w; g->h = h; g->d = d; g->bw = bw; g->bh = bh; g->bd = bd;
    g->brw = bw / 2; g->brh = bh / 2; g->brd = bd / 2;
    g->K = footprint3d(bw, bh, bd, nullptr, nullptr, nullptr);
    if (g->K < 2 || g->K > kMaxK) {
        set_error("unsupported block size");
        return SRED_ERR_ARG;
    }
    return SRED_OK;
}

double java_float_literal(float f)
{
    if (f == 0.0f) return 0.0;
    char buf[64];
    for (int p = 1; p <= 9; p++) {
        snprintf(buf, sizeof buf, "%.*e", p - 1, (double)f);
        if (strtof(buf, nullptr) == f) return strtod(buf, nullptr);
    }
    return (double)f;
}

int make_geometry(int w, int h, int bw, int bh, Geometry *g)
{
    if (w <= 0 || h <= 0) {
        set_error("image dimensions must be positive");
        return SRED_ERR_ARG;
    }
    if (bw < 3 || bh < 3 || bw % 2 == 0 || bh % 2 == 0) {
        set_error("Block dimensions must be odd (e.g., 3x3 or 5x5). Please try again.");
        return SRED_ERR_ARG;
    }
    if (bw > SRED_MAX_BLOCK_SIDE || bh > SRED_MAX_BLOCK_SIDE) {
        set_error("block side larger than SRED_MAX_BLOCK_SIDE");
        return SRED_ERR_ARG;
    }
    if ((int64_t)w * h >= (int64_t)1 << 31) {
        set_error("image larger than 2^31 pixels");
        return SRED_ERR_ARG;
    }
    g->w = w; g->h = h; g->bw = bw; g->bh = bh; g->brw = bw / 2; g->brh = bh / 2;
    g->K = footprint(bw, bh, nullptr, nullptr);
    if (g->K < 2 || g->K > kMaxK) {
        set_error("unsupported block footprint");
        return SRED_ERR_ARG;
    }
    return SRED_OK;
}

int upload_offsets(sred_ctx *ctx, const Geometry &g, int pitch, DevBuf *buf, int *key)
{
    if (g.bd > 1) {  // ellipsoid: dz*(w*h) + dy*pitch + dx, uploaded every time (the key marks 2D offsets)
        std::vector<int> dx, dy, dz;
        footprint3d(g.bw, g.bh, g.bd, &dx, &dy, &dz);
        std::vector<int> off(g.K);
        const int64_t plane = (int64_t)g.w * g.h;
        for (int k = 0; k < g.K; k++) off[k] = (int)(dz[k] * plane + (int64_t)dy[k] * pitch + dx[k]);
        SRED_TRY(buf->ensure(sizeof(int) * g.K));
        SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        SRED_HIP_CHECK(hipMemcpy(buf->ptr, off.data(), sizeof(int) * g.K, hipMemcpyHostToDevice));
        key[0] = key[1] = key[2] = -1;
        return SRED_OK;
    }
    if (key[0] == g.bw && key[1] == g.bh && key[2] == pitch && buf->ptr) return SRED_OK;
    std::vector<int> dx, dy;
    footprint(g.bw, g.bh, &dx, &dy);
    std::vector<int> off(g.K);
    for (int k = 0; k < g.K; k++) off[k] = dy[k] * pitch + dx[k];
    SRED_TRY(buf->ensure(sizeof(int) * g.K));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    SRED_HIP_CHECK(hipMemcpy(buf->ptr, off.data(), sizeof(int) * g.K, hipMemcpyHostToDevice));
    key[0] = g.bw; key[1] = g.bh; key[2] = pitch;
    return SRED_OK;
}

int time_begin(sred_ctx *ctx)
{
    if (!ctx->ev0) {
        SRED_HIP_CHECK(hipEventCreate(&ctx->ev0));
        SRED_HIP_CHECK(hipEventCreate(&ctx->ev1));
    }
    ctx->last_launches = 0;
    ctx->summed_ms = -1.0f;
    SRED_HIP_CHECK(hipEventRecord(ctx->ev0, ctx->stream));
    return SRED_OK;
}

int time_end(sred_ctx *ctx, int64_t units)
{
    SRED_HIP_CHECK(hipEventRecord(ctx->ev1, ctx->stream));
    ctx->last_launches = 1;
    ctx->last_pairs = units;
    return SRED_OK;
}

static int check_ctx(sred_ctx *ctx)
{
    if (!ctx) {
        set_error("null context");
        return SRED_ERR_ARG;
    }
    SRED_HIP_CHECK(hipSetDevice(ctx->device));
    return SRED_OK;
}

static int upload(sred_ctx *ctx, DevBuf &b, const float *h, int64_t n)
{
    SRED_TRY(b.ensure(sizeof(float) * n));
    SRED_HIP_CHECK(hipMemcpyAsync(b.ptr, h, sizeof(float) * n, hipMemcpyHostToDevice, ctx->stream));
    return SRED_OK;
}

static int download(sred_ctx *ctx, float *h, const DevBuf &b, int64_t n)
{
    SRED_HIP_CHECK(hipMemcpyAsync(h, b.ptr, sizeof(float) * n, hipMemcpyDeviceToHost, ctx->stream));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SRED_OK;
}

// Block-repetition host sequence of CLUtils.getBlock*2D on device maps:
// Pearson/SSIM CLUtils.java:565-577, NRMSE :788-841, AbsDiff :938-975.
int block_pipeline(sred_ctx *ctx, const Geometry &g, int metric, const float *d_img, const float *tmpl,
                   float tmpl_mean, float tmpl_std, float c, int normalize, float *d_out)
{
    const int64_t n = g.n();
    const bool need_stats = (c != 0.0f);
    if (need_stats) {
        SRED_TRY(ctx->means.ensure(sizeof(float) * n));
        SRED_TRY(ctx->stds.ensure(sizeof(float) * n));
    }
    // the first normalisation directly follows the stencil (unmasked): the
    // stencil's workgroups supply its min/max partials
    const bool fused = (metric == SRED_ABSDIFF) || ((metric == SRED_PEARSON || metric == SRED_SSIM) && c == 0.0f &&
                                                    normalize);
    int nparts = 0;
    unsigned int *d_parts = nullptr;
    if (fused) {
        SRED_TRY(ctx->mmpart.ensure(sizeof(unsigned int) * 3 * (size_t)block_workgroups(g)));
        d_parts = ctx->mmpart.as<unsigned int>();
    }
    SRED_TRY(block_raw_dev(ctx, g, metric, d_img, need_stats ? ctx->means.as<float>() : nullptr,
                           need_stats ? ctx->stds.as<float>() : nullptr, tmpl, tmpl_mean, tmpl_std, d_out, d_parts,
                           &nparts));
    float *d_mask = nullptr;
    if (need_stats) {
        float nv, np;
        SRED_TRY(noise_var_dev(ctx, d_img, g.w, g.h, &nv));
        SRED_TRY(ctx->mask.ensure(sizeof(float) * n));
        d_mask = ctx->mask.as<float>();
        SRED_TRY(relevance_mask_dev(ctx, g, ctx->stds.as<float>(), nv * c, d_mask, &np));
    }
    if (metric == SRED_PEARSON || metric == SRED_SSIM) {
        if (fused)
            SRED_TRY(normalize_map_from_parts_dev(ctx, g, d_out, d_parts, nparts, 0));
        else {
            if (c > 0.0f) SRED_TRY(apply_mask_dev(ctx, g, d_out, d_mask));
            if (normalize) SRED_TRY(normalize_map_dev(ctx, g, d_out, c > 0.0f ? d_mask : nullptr, 0));
        }
    } else if (metric == SRED_NRMSE) {
        if (c == 0.0f) {
            SRED_TRY(ctx->mask.ensure(sizeof(float) * n));
            d_mask = ctx->mask.as<float>();
            SRED_TRY(interior_ones_mask_dev(ctx, g, d_mask));
        }
        SRED_TRY(apply_mask_dev(ctx, g, d_out, d_mask));
        SRED_TRY(normalize_map_dev(ctx, g, d_out, d_mask, 0));
        SRED_TRY(invert_masked_dev(ctx, g, d_out, d_mask));
        if (normalize) SRED_TRY(normalize_map_dev(ctx, g, d_out, d_mask, 0));
    } else {
        SRED_TRY(normalize_map_from_parts_dev(ctx, g, d_out, d_parts, nparts, 1));  // Float.MIN_VALUE max initialiser
        SRED_TRY(one_minus_dev(ctx, g, d_out));
        if (c > 0.0f) SRED_TRY(apply_mask_dev(ctx, g, d_out, d_mask));
        if (normalize) SRED_TRY(normalize_map_dev(ctx, g, d_out, c > 0.0f ? d_mask : nullptr, 0));
    }
    return SRED_OK;
}

// The time-lapse branch of BlockRepetition2D_.run (BlockRepetition2D_.java:95-137)
// without stabilisation: the stack normalised as a whole (getInputImage3D,
// Utils.java:1257-1260), each frame normalised again (getInputImage2D(float[]),
// Utils.java:760-775) and mapped by the block pipeline, the stacked maps
// normalised as a whole when normalize (getInputImage3D(float[], ...)).
static int block_timelapse_dev(sred_ctx *ctx, int metric, const float *d_stack, int w, int h, int depth,
                               const float *tmpl, int bw, int bh, float tmpl_mean, float tmpl_std, float c,
                               int normalize, float *d_out)
{
    if (depth < 3) {
        set_error("Image must have at least 3 slices. Please try again.");
        return SRED_ERR_ARG;
    }
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    const int64_t n = g.n(), total = n * depth;
    SRED_TRY(ctx->stack_work.ensure(sizeof(float) * total));
    float *work = ctx->stack_work.as<float>();
    SRED_TRY(stack_frames_normalize_dev(ctx, d_stack, work, n, depth));
    if ((metric == SRED_PEARSON || metric == SRED_SSIM) && c == 0.0f && normalize) {
        // the default configuration: all frames in one stencil launch (grid z = frame), whose
        // workgroups supply each frame's min/max partials for its normalisation
        for (int t0 = 0; t0 < depth; t0 += 65535) {
            const int nf = std::min(depth - t0, 65535);
            const int64_t nwg = block_workgroups(g);
            SRED_TRY(ctx->mmpart.ensure(sizeof(unsigned int) * 3 * (size_t)nwg * nf));
            int nparts = 0;
            SRED_TRY(block_raw_dev(ctx, g, metric, work + t0 * n, nullptr, nullptr, tmpl, tmpl_mean, tmpl_std,
                                   d_out + t0 * n, ctx->mmpart.as<unsigned int>(), &nparts, nf));
            SRED_TRY(normalize_frames_from_parts_dev(ctx, g, d_out + t0 * n, nf, ctx->mmpart.as<unsigned int>(),
                                                     nparts, 0));
        }
    } else {
        for (int t = 0; t < depth; t++)
            SRED_TRY(block_pipeline(ctx, g, metric, work + t * n, tmpl, tmpl_mean, tmpl_std, c, normalize,
                                    d_out + t * n));
    }
    if (normalize) SRED_TRY(seg_normalize_dev(ctx, d_out, d_out, total, 1));
    return SRED_OK;
}

}  // namespace sred

using namespace sred;

extern "C" {

int sred_version(void) { return SRED_VERSION; }

const char *sred_last_error(void) { return g_last_error.c_str(); }

int sred_ctx_create(int device, sred_ctx **out)
{
    if (!out) {
        set_error("null output pointer");
        return SRED_ERR_ARG;
    }
    *out = nullptr;
    int ndev = 0;
    hipError_t e = hipGetDeviceCount(&ndev);
    if (e != hipSuccess || ndev == 0) {
        set_error("no HIP device visible");
        return SRED_ERR_NO_DEVICE;
    }
    if (device < 0 || device >= ndev) {
        set_error("device index out of range");
        return SRED_ERR_ARG;
    }
    SRED_HIP_CHECK(hipSetDevice(device));
    // set-up with the device current: the no-allocation entry point's
    // scratch address on this device, and the timing events
    unsigned char *scratch = nullptr;
    SRED_TRY(noalloc_scratch_init(device, &scratch));
    hipEvent_t e0 = nullptr, e1 = nullptr;
    SRED_HIP_CHECK(hipEventCreate(&e0));
    e = hipEventCreate(&e1);
    if (e != hipSuccess) {
        (void)hipEventDestroy(e0);
        set_error(std::string("hipEventCreate: ") + hipGetErrorString(e));
        return SRED_ERR_HIP;
    }
    sred_ctx *c = new sred_ctx();
    c->device = device;
    c->noalloc_scratch = scratch;
    c->ev0 = e0;
    c->ev1 = e1;
    *out = c;
    return SRED_OK;
}

int sred_ctx_destroy(sred_ctx *ctx)
{
    if (!ctx) return SRED_OK;
    (void)hipSetDevice(ctx->device);
    (void)hipStreamSynchronize(ctx->stream);
    DevBuf *bufs[] = {&ctx->img, &ctx->means, &ctx->stds, &ctx->mask, &ctx->S, &ctx->W, &ctx->out,
                      &ctx->tile_var, &ctx->counts, &ctx->valid_idx, &ctx->aux, &ctx->tflags, &ctx->u32, &ctx->uhi,
                      &ctx->ulo, &ctx->partial, &ctx->misc, &ctx->offs, &ctx->offs2, &ctx->tmpl, &ctx->dmap, &ctx->grad,
                      &ctx->wmom, &ctx->wpart, &ctx->symtab, &ctx->batch_part, &ctx->batch_tab, &ctx->pm_rows,
                      &ctx->pm_cols, &ctx->pm_part, &ctx->pm_spec, &ctx->sortbuf, &ctx->mmpart, &ctx->adj_v,
                      &ctx->qt_res, &ctx->qt_nodes, &ctx->qt_leaves, &ctx->qt_seg, &ctx->qt_tmp,
                      &ctx->stack_work, &ctx->seg_mm, &ctx->tmpl3d, &ctx->offs3d, &ctx->adj_f, &ctx->acc64};
    for (DevBuf *b : bufs) b->release();
    for (BatchStore *b : ctx->batch_store) delete b;
    ctx->batch_store.clear();
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->h_fit) (void)hipHostFree(ctx->h_fit);
    if (ctx->h_stage) (void)hipHostFree(ctx->h_stage);
    delete ctx;
    return SRED_OK;
}

int sred_ctx_set_stream(sred_ctx *ctx, void *stream)
{
    SRED_TRY(check_ctx(ctx));
    ctx->stream = static_cast<hipStream_t>(stream);
    return SRED_OK;
}

int sred_ctx_synchronize(sred_ctx *ctx)
{
    SRED_TRY(check_ctx(ctx));
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    return SRED_OK;
}

int sred_ctx_set_chunk_cols(sred_ctx *ctx, int cols)
{
    if (!ctx || cols < 0) {
        set_error("invalid argument");
        return SRED_ERR_ARG;
    }
    ctx->chunk_cols = cols;
    return SRED_OK;
}

int sred_ctx_set_option(sred_ctx *ctx, int option, int value)
{
    if (!ctx) {
        set_error("null context");
        return SRED_ERR_ARG;
    }
    switch (option) {
    case SRED_OPT_CHUNK_COLS:
        if (value < 0) break;
        ctx->chunk_cols = value;
        return SRED_OK;
    case SRED_OPT_SYMMETRIC:
        ctx->symmetric = value != 0;
        return SRED_OK;
    case SRED_OPT_AUTOTUNE:
        ctx->autotune = value != 0;
        return SRED_OK;
    case SRED_OPT_BATCH_PANEL:
        if (value < 0 || value > 2048 || value % 256 != 0) break;
        ctx->batch_panel = value;
        return SRED_OK;
    case SRED_OPT_BATCH_MEMORY_MB:
        if (value < 1) break;
        ctx->batch_mb = value;
        return SRED_OK;
    case SRED_OPT_FIT_TIMING:
        ctx->fit_timing = value != 0;
        return SRED_OK;
    case SRED_OPT_PARTIAL_MEMORY_MB:
        if (value < 1) break;
        ctx->partial_mb = value;
        return SRED_OK;
    default:
        break;
    }
    set_error("unknown option or invalid value");
    return SRED_ERR_ARG;
}

int sred_metric_from_name(const char *name, int block)
{
    if (!name) {
        set_error("null metric name");
        return SRED_ERR_ARG;
    }
    static const char *global_names[4] = {"Pearson's R", "SSIM", "NRMSE", "Abs. diff. of StdDevs"};
    static const char *block_names[4] = {"Pearson's R", "SSIM", "NRMSE (inverted)", "Abs. Diff. of StdDevs."};
    const char **names = block ? block_names : global_names;
    for (int i = 0; i < 4; i++)
        if (strcmp(name, names[i]) == 0) return i;
    set_error(std::string("unknown metric: ") + name);
    return SRED_ERR_UNKNOWN_METRIC;
}

int sred_block_size_2d(int bw, int bh)
{
    if (bw < 3 || bh < 3 || bw > SRED_MAX_BLOCK_SIDE || bh > SRED_MAX_BLOCK_SIDE) {
        set_error("block side out of range");
        return SRED_ERR_ARG;
    }
    return footprint(bw, bh, nullptr, nullptr);
}

int sred_normalize_input_2d(float *img, int64_t n)
{
#pragma clang fp contract(off)
    if (!img || n <= 0) {
        set_error("invalid argument");
        return SRED_ERR_ARG;
    }
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int64_t i = 0; i < n; i++) {
        mn = std::min(mn, img[i]);
        mx = std::max(mx, img[i]);
    }
    for (int64_t i = 0; i < n; i++) img[i] = (img[i] - mn) / (mx - mn + kEps);
    return SRED_OK;
}

int sred_reference_block_2d(const float *block, int bw, int bh, float *px, float *mean, float *std_)
{
#pragma clang fp contract(off)
    if (!block || !px || !mean || !std_) {
        set_error("invalid argument");
        return SRED_ERR_ARG;
    }
    if (bw % 2 == 0 || bh % 2 == 0 || bw < 3 || bh < 3 || bw > SRED_MAX_BLOCK_SIDE || bh > SRED_MAX_BLOCK_SIDE) {
        set_error("Block dimensions must be odd (e.g., 3x3 or 5x5). Please try again.");
        return SRED_ERR_ARG;
    }
    std::vector<int> dx, dy;
    int K = footprint(bw, bh, &dx, &dy);
    const int brw = bw / 2, brh = bh / 2;
    for (int k = 0; k < K; k++) px[k] = block[(dy[k] + brh) * bw + dx[k] + brw];
    float bmin = FLT_MAX, bmax = -FLT_MAX;
    for (int k = 0; k < K; k++) { bmin = std::min(bmin, px[k]); bmax = std::max(bmax, px[k]); }
    float m = 0.0f;
    for (int k = 0; k < K; k++) {
        px[k] = (px[k] - bmin) / (bmax - bmin + kEps);
        m += px[k];
    }
    m /= (float)K;
    for (int k = 0; k < K; k++) px[k] = px[k] - m;
    float s = 0.0f;
    for (int k = 0; k < K; k++) s += (px[k] - m) * (px[k] - m);  // Utils.java:632-634 subtracts the mean again
    s = (float)std::sqrt((double)(s / (float)(K - 1)));
    *mean = m;
    *std_ = s;
    return K;
}

int sred_noise_var_2d(const float *img, int w, int h, float *noise_var)
{
#pragma clang fp contract(off)
    if (!img || !noise_var || w <= 0 || h <= 0) {
        set_error("invalid argument");
        return SRED_ERR_ARG;
    }
    int tbw = 8, tbh = 8;
    if ((int64_t)w * h > 352 * 288) { tbw = 16; tbh = 16; }
    int nbx = w / tbw, nby = h / tbh, nb = nbx * nby;
    if (nb <= 0) {
        set_error("image smaller than one noise tile (the reference indexes an empty variance array)");
        return SRED_ERR_ARG;
    }
    std::vector<float> lv(nb);
    for (int by = 0; by < nby; by++)
        for (int bx = 0; bx < nbx; bx++) {
            float m = 0.0f, sq = 0.0f;
            for (int y = by * tbh; y < (by + 1) * tbh; y++)
                for (int x = bx * tbw; x < (bx + 1) * tbw; x++) {
                    float v = img[(int64_t)y * w + x];
                    if (!std::isfinite(v)) {
                        set_error("Input contains invalid (NaN or Infinite) values.");
                        return SRED_ERR_NONFINITE;
                    }
                    m += v;
                    sq += v * v;
                }
            m /= (float)(tbw * tbh);
            lv[by * nbx + bx] = sq / (float)(tbw * tbh) - m * m;
        }
    std::sort(lv.begin(), lv.end());
    int nvars = (int)(0.03f * (float)nb + 1.0f);
    float nv = 0.0f;
    for (int i = 0; i < nvars && i < nb; i++) nv += lv[i];
    nv = std::fabs(nv / (float)nvars);
    nv = (1.0f + 0.001f * (nv - 40.0f)) * nv;
    *noise_var = nv;
    return SRED_OK;
}

int sred_relevance_mask_2d(const float *stds, int w, int h, int brw, int brh, float noise_var, float c,
                           float *mask, float *threshold, float *n_pixels)
{
#pragma clang fp contract(off)
    if (!stds || !mask || w <= 0 || h <= 0 || brw < 0 || brh < 0) {
        set_error("invalid argument");
        return SRED_ERR_ARG;
    }
    float thr = noise_var * c;
    float np = 0.0f;
    for (int y = 0; y < h; y++)
        for (int x = 0; x < w; x++) {
            int64_t i = (int64_t)y * w + x;
            float m = 0.0f;
            if (x >= brw && x < w - brw && y >= brh && y < h - brh) m = (stds[i] * stds[i] <= thr) ? 0.0f : 1.0f;
            mask[i] = m;
            if (m > 0.0f) np += 1.0f;
        }
    if (threshold) *threshold = thr;
    if (n_pixels) *n_pixels = np;
    return SRED_OK;
}

int sred_apply_mask_2d(float *map, int w, int h, const float *mask)
{
    if (!map || !mask || w <= 0 || h <= 0) {
        set_error("invalid argument");
        return SRED_ERR_ARG;
    }
    for (int64_t i = 0; i < (int64_t)w * h; i++) map[i] *= mask[i];
    return SRED_OK;
}

int sred_normalize_map_2d(float *map, int w, int h, int brw, int brh, const float *mask)
{
#pragma clang fp contract(off)
    if (!map || w <= 0 || h <= 0) {
        set_error("invalid argument");
        return SRED_ERR_ARG;
    }
    float mn = FLT_MAX, mx = -FLT_MAX;
    for (int y = brh; y < h - brh; y++)
        for (int x = brw; x < w - brw; x++) {
            int64_t i = (int64_t)y * w + x;
            if (mask) {
                if (mask[i] != 0.0f && mask[i] != 1.0f) {
                    set_error("Relevance mask must be binary (0.0f or 1.0f)");
                    return SRED_ERR_NONBINARY_MASK;
                }
                if (!(mask[i] > 0.0f)) continue;
            }
            mn = std::min(map[i], mn);
            mx = std::max(map[i], mx);
        }
    for (int y = brh; y < h - brh; y++)
        for (int x = brw; x < w - brw; x++) {
            int64_t i = (int64_t)y * w + x;
            if (mask && !(mask[i] > 0.0f)) continue;
            map[i] = (map[i] - mn) / (mx - mn + kEps);
        }
    return SRED_OK;
}

int sred_local_stats_2d(sred_ctx *ctx, const float *img, int w, int h, int bw, int bh, float *means, float *stds)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!img || !means || !stds) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, img, g.n()));
    SRED_TRY(ctx->means.ensure(sizeof(float) * g.n()));
    SRED_TRY(ctx->stds.ensure(sizeof(float) * g.n()));
    SRED_TRY(launch_local_stats(ctx, g, ctx->img.as<float>(), ctx->means.as<float>(), ctx->stds.as<float>()));
    SRED_HIP_CHECK(hipMemcpyAsync(means, ctx->means.ptr, sizeof(float) * g.n(), hipMemcpyDeviceToHost, ctx->stream));
    SRED_TRY(download(ctx, stds, ctx->stds, g.n()));
    return SRED_OK;
}

int sred_global_accum_2d(sred_ctx *ctx, int metric, int precision, const float *img, const float *means,
                         const float *stds, int w, int h, int bw, int bh, float filter_param, float threshold,
                         float *S, float *Wsum)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!img || !means || !stds || !S || !Wsum) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    if (precision != SRED_FP32 && precision != SRED_BF16X3) {
        set_error("unknown precision");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, img, g.n()));
    SRED_TRY(upload(ctx, ctx->means, means, g.n()));
    SRED_TRY(upload(ctx, ctx->stds, stds, g.n()));
    SRED_TRY(ctx->S.ensure(sizeof(float) * g.n()));
    SRED_TRY(ctx->W.ensure(sizeof(float) * g.n()));
    GlobalParams p;
    p.metric = metric;
    p.precision = precision;
    p.filter_param = filter_param;
    p.threshold = threshold;
    SRED_TRY(global_accum_dev(ctx, g, p, ctx->img.as<float>(), ctx->means.as<float>(), ctx->stds.as<float>(),
                              ctx->S.as<float>(), ctx->W.as<float>()));
    SRED_HIP_CHECK(hipMemcpyAsync(S, ctx->S.ptr, sizeof(float) * g.n(), hipMemcpyDeviceToHost, ctx->stream));
    SRED_TRY(download(ctx, Wsum, ctx->W, g.n()));
    return SRED_OK;
}

int sred_global_repetition_2d(sred_ctx *ctx, int metric, int precision, const float *img, const float *means,
                              const float *stds, int w, int h, int bw, int bh, const float *mask,
                              float relevance_constant, float threshold, float noise_var, float n_pixels,
                              int normalize, float *out)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!img || !means || !stds || !mask || !out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    if (metric < 0 || metric > 3) {
        set_error("unknown metric");
        return SRED_ERR_UNKNOWN_METRIC;
    }
    if (precision != SRED_FP32 && precision != SRED_BF16X3) {
        set_error("unknown precision");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, img, g.n()));
    SRED_TRY(upload(ctx, ctx->means, means, g.n()));
    SRED_TRY(upload(ctx, ctx->stds, stds, g.n()));
    SRED_TRY(upload(ctx, ctx->mask, mask, g.n()));
    if (normalize) {
        int ok;
        SRED_TRY(mask_is_binary_dev(ctx, g, ctx->mask.as<float>(), &ok));
        if (!ok) {
            set_error("Relevance mask must be binary (0.0f or 1.0f)");
            return SRED_ERR_NONBINARY_MASK;
        }
    }
    SRED_TRY(ctx->S.ensure(sizeof(float) * g.n()));
    SRED_TRY(ctx->W.ensure(sizeof(float) * g.n()));
    SRED_TRY(ctx->out.ensure(sizeof(float) * g.n()));
    GlobalParams p;
    p.metric = metric;
    p.precision = precision;
    p.filter_param = noise_var;
    p.threshold = threshold;
    SRED_TRY(global_accum_dev(ctx, g, p, ctx->img.as<float>(), ctx->means.as<float>(), ctx->stds.as<float>(),
                              ctx->S.as<float>(), ctx->W.as<float>()));
    SRED_TRY(global_finalize_dev(ctx, g, metric, ctx->S.as<float>(), ctx->W.as<float>(), ctx->mask.as<float>(),
                                 n_pixels, relevance_constant, normalize, ctx->out.as<float>()));
    SRED_TRY(download(ctx, out, ctx->out, g.n()));
    return SRED_OK;
}

int sred_block_repetition_2d(sred_ctx *ctx, int metric, const float *img, int w, int h, const float *tmpl, int bw,
                             int bh, float tmpl_mean, float tmpl_std, float relevance_constant, int normalize,
                             float *out)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!img || !tmpl || !out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    if (metric < 0 || metric > 3) {
        set_error("unknown metric");
        return SRED_ERR_UNKNOWN_METRIC;
    }
    SRED_TRY(upload(ctx, ctx->img, img, g.n()));
    SRED_TRY(ctx->out.ensure(sizeof(float) * g.n()));
    SRED_TRY(block_pipeline(ctx, g, metric, ctx->img.as<float>(), tmpl, tmpl_mean, tmpl_std, relevance_constant,
                            normalize, ctx->out.as<float>()));
    SRED_TRY(download(ctx, out, ctx->out, g.n()));
    return SRED_OK;
}

int sred_noise_var_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, float *noise_var)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_img || !noise_var || w <= 0 || h <= 0) {
        set_error("invalid argument");
        return SRED_ERR_ARG;
    }
    return noise_var_dev(ctx, d_img, w, h, noise_var);
}

int sred_global_map_2d_dev(sred_ctx *ctx, int metric, int precision, const float *d_img, int w, int h, int bw,
                           int bh, float relevance_constant, int normalize, float *d_out, float *d_means,
                           float *d_stds, float *d_mask, float *d_S, float *d_Wsum)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!d_img || !d_out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    if (metric < 0 || metric > 3) {
        set_error("unknown metric");
        return SRED_ERR_UNKNOWN_METRIC;
    }
    if (precision != SRED_FP32 && precision != SRED_BF16X3) {
        set_error("unknown precision");
        return SRED_ERR_ARG;
    }
    const int64_t n = g.n();
    if (!d_means) { SRED_TRY(ctx->means.ensure(sizeof(float) * n)); d_means = ctx->means.as<float>(); }
    if (!d_stds) { SRED_TRY(ctx->stds.ensure(sizeof(float) * n)); d_stds = ctx->stds.as<float>(); }
    if (!d_mask) { SRED_TRY(ctx->mask.ensure(sizeof(float) * n)); d_mask = ctx->mask.as<float>(); }
    if (!d_S) { SRED_TRY(ctx->S.ensure(sizeof(float) * n)); d_S = ctx->S.as<float>(); }
    if (!d_Wsum) { SRED_TRY(ctx->W.ensure(sizeof(float) * n)); d_Wsum = ctx->W.as<float>(); }
    // GlobalRepetition2D_.java:124-140
    SRED_TRY(launch_local_stats(ctx, g, d_img, d_means, d_stds));
    float nv, np;
    SRED_TRY(noise_var_dev(ctx, d_img, g.w, g.h, &nv));
    const float thr = nv * relevance_constant;
    SRED_TRY(relevance_mask_dev(ctx, g, d_stds, thr, d_mask, &np));
    GlobalParams p;
    p.metric = metric;
    p.precision = precision;
    p.filter_param = nv;
    p.threshold = thr;
    SRED_TRY(global_accum_dev(ctx, g, p, d_img, d_means, d_stds, d_S, d_Wsum));
    SRED_TRY(global_finalize_dev(ctx, g, metric, d_S, d_Wsum, d_mask, np, relevance_constant, normalize, d_out));
    return SRED_OK;
}

int sred_block_map_2d_dev(sred_ctx *ctx, int metric, const float *d_img, int w, int h, const float *tmpl, int bw,
                          int bh, float tmpl_mean, float tmpl_std, float relevance_constant, int normalize,
                          float *d_out)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!d_img || !tmpl || !d_out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    if (metric < 0 || metric > 3) {
        set_error("unknown metric");
        return SRED_ERR_UNKNOWN_METRIC;
    }
    return block_pipeline(ctx, g, metric, d_img, tmpl, tmpl_mean, tmpl_std, relevance_constant, normalize, d_out);
}

int sred_block_hu_2d(sred_ctx *ctx, const float *img, int w, int h, const float *tmpl, int bw, int bh, float *out)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!img || !tmpl || !out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, img, g.n()));
    SRED_TRY(ctx->out.ensure(sizeof(float) * g.n()));
    SRED_TRY(block_hu_dev(ctx, g, ctx->img.as<float>(), tmpl, ctx->out.as<float>()));
    SRED_TRY(download(ctx, out, ctx->out, g.n()));
    return SRED_OK;
}

int sred_block_hu_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, const float *tmpl, int bw, int bh,
                         float *d_out)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!d_img || !tmpl || !d_out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    return block_hu_dev(ctx, g, d_img, tmpl, d_out);
}

// All-pairs vector maps: host-buffer entries upload img/means/stds, run the
// device path and download the map.
static int pair_map_host(sred_ctx *ctx, int which, const float *img, const float *means, const float *stds, int w,
                         int h, int bw, int bh, const int32_t *rows, int64_t nrows, float *out)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!img || !means || (which != 2 && !stds) || !out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, img, g.n()));
    SRED_TRY(upload(ctx, ctx->means, means, g.n()));
    if (which != 2) SRED_TRY(upload(ctx, ctx->stds, stds, g.n()));
    SRED_TRY(ctx->out.ensure(sizeof(float) * g.n()));
    float *d_out = ctx->out.as<float>();
    if (which == 0)
        SRED_TRY(global_hu_dev(ctx, g, ctx->img.as<float>(), ctx->means.as<float>(), ctx->stds.as<float>(), rows,
                               nrows, d_out));
    else if (which == 1)
        SRED_TRY(phase_map_dev(ctx, g, ctx->img.as<float>(), ctx->means.as<float>(), ctx->stds.as<float>(), rows,
                               nrows, d_out));
    else
        SRED_TRY(hu_invariants_dev(ctx, g, ctx->img.as<float>(), ctx->means.as<float>(), d_out));
    return download(ctx, out, ctx->out, g.n());
}

static int pair_map_dev(sred_ctx *ctx, int which, const float *d_img, const float *d_means, const float *d_stds,
                        int w, int h, int bw, int bh, const int32_t *rows, int64_t nrows, float *d_out)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!d_img || !d_means || !d_stds || !d_out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    if (which == 0) return global_hu_dev(ctx, g, d_img, d_means, d_stds, rows, nrows, d_out);
    return phase_map_dev(ctx, g, d_img, d_means, d_stds, rows, nrows, d_out);
}

int sred_hu_invariants_2d(sred_ctx *ctx, const float *img, const float *means, int w, int h, int bw, int bh,
                          float *out)
{
    return pair_map_host(ctx, 2, img, means, nullptr, w, h, bw, bh, nullptr, 0, out);
}

int sred_global_hu_2d(sred_ctx *ctx, const float *img, const float *means, const float *stds, int w, int h, int bw,
                      int bh, const int32_t *rows, int64_t nrows, float *out)
{
    return pair_map_host(ctx, 0, img, means, stds, w, h, bw, bh, rows, nrows, out);
}

int sred_global_hu_2d_dev(sred_ctx *ctx, const float *d_img, const float *d_means, const float *d_stds, int w, int h,
                          int bw, int bh, const int32_t *rows, int64_t nrows, float *d_out)
{
    return pair_map_dev(ctx, 0, d_img, d_means, d_stds, w, h, bw, bh, rows, nrows, d_out);
}

int sred_phase_correlation_2d(sred_ctx *ctx, const float *img, const float *means, const float *stds, int w, int h,
                              int bw, int bh, const int32_t *rows, int64_t nrows, float *out)
{
    return pair_map_host(ctx, 1, img, means, stds, w, h, bw, bh, rows, nrows, out);
}

int sred_phase_correlation_2d_dev(sred_ctx *ctx, const float *d_img, const float *d_means, const float *d_stds,
                                  int w, int h, int bw, int bh, const int32_t *rows, int64_t nrows, float *d_out)
{
    return pair_map_dev(ctx, 1, d_img, d_means, d_stds, w, h, bw, bh, rows, nrows, d_out);
}

int sred_gat_2d(sred_ctx *ctx, float *img, int64_t n, double gain, double sigma, double offset)
{
    SRED_TRY(check_ctx(ctx));
    if (!img || n < 0) {
        set_error("null buffer or negative size");
        return SRED_ERR_ARG;
    }
    if (n == 0) return SRED_OK;
    SRED_TRY(upload(ctx, ctx->img, img, n));
    SRED_TRY(gat_dev(ctx, ctx->img.as<float>(), n, gain, sigma, offset));
    return download(ctx, img, ctx->img, n);
}

int sred_gat_2d_dev(sred_ctx *ctx, float *d_img, int64_t n, double gain, double sigma, double offset)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_img || n < 0) {
        set_error("null buffer or negative size");
        return SRED_ERR_ARG;
    }
    return gat_dev(ctx, d_img, n, gain, sigma, offset);
}

int sred_gat_objective_2d(sred_ctx *ctx, const float *img, int w, int h, double gain, double sigma, double offset,
                          double *error)
{
    SRED_TRY(check_ctx(ctx));
    if (!img || !error || w <= 0 || h <= 0) {
        set_error("null buffer or empty image");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, img, (int64_t)w * h));
    return gat_objective_dev(ctx, ctx->img.as<float>(), w, h, gain, sigma, offset, error);
}

int sred_gat_minimize_2d(sred_ctx *ctx, const float *img, int w, int h, double gain0, double sigma0, double offset0,
                         int max_iter, double *params, double *error)
{
    SRED_TRY(check_ctx(ctx));
    if (!img || !params || w <= 0 || h <= 0 || max_iter < 0) {
        set_error("null buffer, empty image or negative iteration count");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, img, (int64_t)w * h));
    return gat_minimize_dev(ctx, ctx->img.as<float>(), w, h, gain0, sigma0, offset0, max_iter, params, error);
}

int sred_gat_minimize_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, double gain0, double sigma0,
                             double offset0, int max_iter, double *params, double *error)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_img || !params || w <= 0 || h <= 0 || max_iter < 0) {
        set_error("null buffer, empty image or negative iteration count");
        return SRED_ERR_ARG;
    }
    return gat_minimize_dev(ctx, d_img, w, h, gain0, sigma0, offset0, max_iter, params, error);
}

int sred_relevance_mask_stack_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh, float lower,
                                    float upper, float step, int cap, float *d_masks, float *constants, int *n_masks)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_img || !n_masks || cap < 0 || (cap > 0 && !d_masks)) {
        set_error("null buffer or negative capacity");
        return SRED_ERR_ARG;
    }
    return relevance_mask_stack_dev(ctx, d_img, w, h, bw, bh, lower, upper, step, cap, d_masks, constants, n_masks);
}

int sred_relevance_mask_stack_2d(sred_ctx *ctx, const float *img, int w, int h, int bw, int bh, float lower,
                                float upper, float step, int cap, float *masks, float *constants, int *n_masks)
{
    SRED_TRY(check_ctx(ctx));
    if (!img || !n_masks || cap < 0 || (cap > 0 && !masks) || w <= 0 || h <= 0) {
        set_error("null buffer, negative capacity or empty image");
        return SRED_ERR_ARG;
    }
    const int64_t n = (int64_t)w * h;
    SRED_TRY(upload(ctx, ctx->img, img, n));
    SRED_TRY(ctx->stack_work.ensure(sizeof(float) * (size_t)std::max(cap, 1) * n));
    SRED_TRY(relevance_mask_stack_dev(ctx, ctx->img.as<float>(), w, h, bw, bh, lower, upper, step, cap,
                                      ctx->stack_work.as<float>(), constants, n_masks));
    return download(ctx, masks, ctx->stack_work, (int64_t)*n_masks * n);
}

int sred_gat_objective_3d(sred_ctx *ctx, const float *vol, int w, int h, int d, double gain, double sigma,
                          double offset, double *error)
{
    SRED_TRY(check_ctx(ctx));
    if (!vol || !error || w <= 0 || h <= 0 || d <= 0 || (int64_t)w * h * d >= (1LL << 31)) {
        set_error("null buffer, empty volume or more than 2^31 voxels");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, vol, (int64_t)w * h * d));
    return gat_objective3d_dev(ctx, ctx->img.as<float>(), w, h, d, gain, sigma, offset, error);
}

int sred_gat_minimize_3d(sred_ctx *ctx, const float *vol, int w, int h, int d, double gain0, double sigma0,
                         double offset0, int max_iter, double *params, double *error)
{
    SRED_TRY(check_ctx(ctx));
    if (!vol || !params || w <= 0 || h <= 0 || d <= 0 || (int64_t)w * h * d >= (1LL << 31) || max_iter < 0) {
        set_error("null buffer, empty volume, more than 2^31 voxels or negative iteration count");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, vol, (int64_t)w * h * d));
    return gat_minimize3d_dev(ctx, ctx->img.as<float>(), w, h, d, gain0, sigma0, offset0, max_iter, params, error);
}

int sred_gat_minimize_3d_dev(sred_ctx *ctx, const float *d_vol, int w, int h, int d, double gain0, double sigma0,
                             double offset0, int max_iter, double *params, double *error)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_vol || !params || w <= 0 || h <= 0 || d <= 0 || (int64_t)w * h * d >= (1LL << 31) || max_iter < 0) {
        set_error("null buffer, empty volume, more than 2^31 voxels or negative iteration count");
        return SRED_ERR_ARG;
    }
    return gat_minimize3d_dev(ctx, d_vol, w, h, d, gain0, sigma0, offset0, max_iter, params, error);
}

int sred_block_timelapse_2d_dev(sred_ctx *ctx, int metric, const float *d_stack, int w, int h, int depth,
                                const float *tmpl, int bw, int bh, float tmpl_mean, float tmpl_std,
                                float relevance_constant, int normalize, float *d_out)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_stack || !tmpl || !d_out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    if (metric < SRED_PEARSON || metric > SRED_ABSDIFF) {
        set_error("unknown metric");
        return SRED_ERR_UNKNOWN_METRIC;
    }
    return block_timelapse_dev(ctx, metric, d_stack, w, h, depth, tmpl, bw, bh, tmpl_mean, tmpl_std,
                               relevance_constant, normalize, d_out);
}

int sred_block_timelapse_2d(sred_ctx *ctx, int metric, const float *stack, int w, int h, int depth,
                            const float *tmpl, int bw, int bh, float tmpl_mean, float tmpl_std,
                            float relevance_constant, int normalize, float *out)
{
    SRED_TRY(check_ctx(ctx));
    if (!stack || !tmpl || !out || w <= 0 || h <= 0 || depth <= 0) {
        set_error("null buffer or empty stack");
        return SRED_ERR_ARG;
    }
    if (metric < SRED_PEARSON || metric > SRED_ABSDIFF) {
        set_error("unknown metric");
        return SRED_ERR_UNKNOWN_METRIC;
    }
    const int64_t total = (int64_t)w * h * depth;
    SRED_TRY(upload(ctx, ctx->img, stack, total));
    SRED_TRY(ctx->out.ensure(sizeof(float) * total));
    SRED_TRY(block_timelapse_dev(ctx, metric, ctx->img.as<float>(), w, h, depth, tmpl, bw, bh, tmpl_mean, tmpl_std,
                                 relevance_constant, normalize, ctx->out.as<float>()));
    return download(ctx, out, ctx->out, total);
}

int sred_global_repetition_3d_dev(sred_ctx *ctx, int metric, int precision, const float *d_img, int w, int h, int d,
                                  int bw, int bh, int bd, float relevance_constant, int normalize, float *d_out,
                                  float *d_means, float *d_stds, float *d_mask, float *d_S, float *d_Wsum)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry3d(w, h, d, bw, bh, bd, &g));
    if (!d_img || !d_out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    if (metric < SRED_PEARSON || metric > SRED_ABSDIFF) {
        set_error("unknown metric");
        return SRED_ERR_UNKNOWN_METRIC;
    }
    if (precision != SRED_FP32 && precision != SRED_BF16X3) {
        set_error("unknown precision");
        return SRED_ERR_ARG;
    }
    const int64_t n = g.n();
    if (!d_means) { SRED_TRY(ctx->means.ensure(sizeof(float) * n)); d_means = ctx->means.as<float>(); }
    if (!d_stds) { SRED_TRY(ctx->stds.ensure(sizeof(float) * n)); d_stds = ctx->stds.as<float>(); }
    if (!d_mask) { SRED_TRY(ctx->mask.ensure(sizeof(float) * n)); d_mask = ctx->mask.as<float>(); }
    if (!d_S) { SRED_TRY(ctx->S.ensure(sizeof(float) * n)); d_S = ctx->S.as<float>(); }
    if (!d_Wsum) { SRED_TRY(ctx->W.ensure(sizeof(float) * n)); d_Wsum = ctx->W.as<float>(); }
    return global_map3d_dev(ctx, g, metric, precision, d_img, relevance_constant, normalize, d_out, d_means, d_stds,
                            d_mask, d_S, d_Wsum);
}

int sred_global_repetition_3d(sred_ctx *ctx, int metric, int precision, const float *img, int w, int h, int d, int bw,
                              int bh, int bd, float relevance_constant, int normalize, float *out)
{
    SRED_TRY(check_ctx(ctx));
    if (!img || !out || w <= 0 || h <= 0 || d <= 0) {
        set_error("null buffer or empty volume");
        return SRED_ERR_ARG;
    }
    const int64_t n = (int64_t)w * h * d;
    SRED_TRY(upload(ctx, ctx->img, img, n));
    SRED_TRY(ctx->out.ensure(sizeof(float) * n));
    SRED_TRY(sred_global_repetition_3d_dev(ctx, metric, precision, ctx->img.as<float>(), w, h, d, bw, bh, bd,
                                           relevance_constant, normalize, ctx->out.as<float>(), nullptr, nullptr,
                                           nullptr, nullptr, nullptr));
    return download(ctx, out, ctx->out, n);
}

int sred_local_stats_3d(sred_ctx *ctx, const float *img, int w, int h, int d, int bw, int bh, int bd, float *means,
                        float *stds)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry3d(w, h, d, bw, bh, bd, &g));
    if (!img || !means || !stds) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    const int64_t n = g.n();
    SRED_TRY(upload(ctx, ctx->img, img, n));
    SRED_TRY(ctx->means.ensure(sizeof(float) * n));
    SRED_TRY(ctx->stds.ensure(sizeof(float) * n));
    SRED_TRY(local_stats3d_dev(ctx, g, ctx->img.as<float>(), ctx->means.as<float>(), ctx->stds.as<float>()));
    SRED_TRY(download(ctx, means, ctx->means, n));
    return download(ctx, stds, ctx->stds, n);
}

int sred_reference_block_3d(const float *block, int bw, int bh, int bd, float *px, float *mean, float *std_)
{
#pragma clang fp contract(off)
    if (!block || !px || !mean || !std_) {
        set_error("invalid argument");
        return SRED_ERR_ARG;
    }
    if (bw % 2 == 0 || bh % 2 == 0 || bd % 2 == 0 || bw < 3 || bh < 3 || bd < 3 || bw > SRED_MAX_BLOCK_SIDE ||
        bh > SRED_MAX_BLOCK_SIDE || bd > SRED_MAX_BLOCK_SIDE) {
        set_error("Block dimensions must be odd (e.g., 3x3x3 or 5x5x5), with at least 3 slices. Please try again.");
        return SRED_ERR_ARG;
    }
    std::vector<int> dx, dy, dz;
    const int K = footprint3d(bw, bh, bd, &dx, &dy, &dz);
    if (K < 2 || K > kMaxK) {
        set_error("unsupported block size");
        return SRED_ERR_ARG;
    }
    const int brw = bw / 2, brh = bh / 2, brd = bd / 2;
    for (int k = 0; k < K; k++) px[k] = block[((dz[k] + brd) * bh + dy[k] + brh) * bw + dx[k] + brw];
    // Utils.normalizeArray, mean subtraction, Utils.normalizeArray (Utils.java:1118-1135)
    for (int pass = 0; pass < 2; pass++) {
        float bmin = FLT_MAX, bmax = -FLT_MAX;
        for (int k = 0; k < K; k++) { bmin = std::min(px[k], bmin); bmax = std::max(px[k], bmax); }
        for (int k = 0; k < K; k++) px[k] = (px[k] - bmin) / (bmax - bmin + kEps);
        if (pass == 0) {
            float m = 0.0f;
            for (int k = 0; k < K; k++) m += px[k];
            m /= (float)K;
            for (int k = 0; k < K; k++) px[k] = px[k] - m;
        }
    }
    float m = 0.0f;
    for (int k = 0; k < K; k++) m += px[k];
    m /= (float)K;
    float s = 0.0f;
    for (int k = 0; k < K; k++) s += (px[k] - m) * (px[k] - m);
    s = (float)std::sqrt((double)(s / (float)(K - 1)));
    *mean = m;
    *std_ = s;
    return K;
}

int sred_block_repetition_3d_dev(sred_ctx *ctx, int metric, const float *d_img, int w, int h, int d,
                                 const float *tmpl, int bw, int bh, int bd, float tmpl_mean, float tmpl_std,
                                 float relevance_constant, int normalize, float *d_out, float *d_means,
                                 float *d_stds, float *d_mask)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry3d(w, h, d, bw, bh, bd, &g));
    if (!d_img || !tmpl || !d_out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    if (bw > w || bh > h || bd > d) {  // BlockRepetition3D_.java:97-100
        set_error("Block dimensions can't be larger than image dimensions.");
        return SRED_ERR_ARG;
    }
    const int64_t n = g.n();
    if (!d_means) { SRED_TRY(ctx->means.ensure(sizeof(float) * n)); d_means = ctx->means.as<float>(); }
    if (!d_stds) { SRED_TRY(ctx->stds.ensure(sizeof(float) * n)); d_stds = ctx->stds.as<float>(); }
    if (!d_mask && relevance_constant > 0.0f) {
        SRED_TRY(ctx->mask.ensure(sizeof(float) * n));
        d_mask = ctx->mask.as<float>();
    }
    return block_map3d_dev(ctx, g, metric, d_img, tmpl, tmpl_mean, tmpl_std, relevance_constant, normalize, d_out,
                           d_means, d_stds, d_mask);
}

int sred_block_repetition_3d(sred_ctx *ctx, int metric, const float *img, int w, int h, int d, const float *tmpl,
                             int bw, int bh, int bd, float tmpl_mean, float tmpl_std, float relevance_constant,
                             int normalize, float *out)
{
    SRED_TRY(check_ctx(ctx));
    if (!img || !out || w <= 0 || h <= 0 || d <= 0) {
        set_error("null buffer or empty volume");
        return SRED_ERR_ARG;
    }
    const int64_t n = (int64_t)w * h * d;
    SRED_TRY(upload(ctx, ctx->img, img, n));
    SRED_TRY(ctx->out.ensure(sizeof(float) * n));
    SRED_TRY(sred_block_repetition_3d_dev(ctx, metric, ctx->img.as<float>(), w, h, d, tmpl, bw, bh, bd, tmpl_mean,
                                          tmpl_std, relevance_constant, normalize, ctx->out.as<float>(), nullptr,
                                          nullptr, nullptr));
    return download(ctx, out, ctx->out, n);
}

int sred_block_size_3d(int bw, int bh, int bd)
{
    if (bw < 3 || bh < 3 || bd < 3 || bw % 2 == 0 || bh % 2 == 0 || bd % 2 == 0) {
        set_error("Block dimensions must be odd, and the block depth at least 3. Please try again.");
        return SRED_ERR_ARG;
    }
    return footprint3d(bw, bh, bd, nullptr, nullptr, nullptr);
}

int sred_tool_prep_2d_dev(sred_ctx *ctx, float *d_img, int w, int h, int max_iter, double *params)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_img || w < 6 || h < 6 || max_iter < 0) {
        set_error("null buffer, an image below 6 x 6 pixels or a negative iteration count");
        return SRED_ERR_ARG;
    }
    return tool_prep_dev(ctx, d_img, w, h, max_iter, params);
}

int sred_tool_prep_2d(sred_ctx *ctx, float *img, int w, int h, int max_iter, double *params)
{
    SRED_TRY(check_ctx(ctx));
    if (!img || w < 6 || h < 6 || max_iter < 0) {
        set_error("null buffer, an image below 6 x 6 pixels or a negative iteration count");
        return SRED_ERR_ARG;
    }
    const int64_t n = (int64_t)w * h;
    SRED_TRY(upload(ctx, ctx->img, img, n));
    SRED_TRY(tool_prep_dev(ctx, ctx->img.as<float>(), w, h, max_iter, params));
    return download(ctx, img, ctx->img, n);
}

int sred_relevance_mask_tool_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh,
                                    float filter_constant, float *d_mask, float *noise_var, float *threshold)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_img || !d_mask) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    return relevance_mask_tool_dev(ctx, d_img, w, h, bw, bh, filter_constant, d_mask, noise_var, threshold);
}

int sred_relevance_mask_tool_2d(sred_ctx *ctx, const float *img, int w, int h, int bw, int bh, float filter_constant,
                                float *mask, float *noise_var, float *threshold)
{
    SRED_TRY(check_ctx(ctx));
    if (!img || !mask || w <= 0 || h <= 0) {
        set_error("null buffer or empty image");
        return SRED_ERR_ARG;
    }
    const int64_t n = (int64_t)w * h;
    SRED_TRY(upload(ctx, ctx->img, img, n));
    SRED_TRY(ctx->mask.ensure(sizeof(float) * n));
    SRED_TRY(relevance_mask_tool_dev(ctx, ctx->img.as<float>(), w, h, bw, bh, filter_constant, ctx->mask.as<float>(),
                                     noise_var, threshold));
    return download(ctx, mask, ctx->mask, n);
}

int sred_optimise_relevance_mask_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh, float step,
                                        int64_t max_steps, float *all_ones, float *all_zeros)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_img || !all_ones || !all_zeros) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    return optimise_relevance_mask_dev(ctx, d_img, w, h, bw, bh, step, max_steps, all_ones, all_zeros);
}

int sred_optimise_relevance_mask_2d(sred_ctx *ctx, const float *img, int w, int h, int bw, int bh, float step,
                                    int64_t max_steps, float *all_ones, float *all_zeros)
{
    SRED_TRY(check_ctx(ctx));
    if (!img || !all_ones || !all_zeros || w <= 0 || h <= 0) {
        set_error("null buffer or empty image");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, img, (int64_t)w * h));
    return optimise_relevance_mask_dev(ctx, ctx->img.as<float>(), w, h, bw, bh, step, max_steps, all_ones, all_zeros);
}

int sred_f_quantile(double d1, double d2, double p, double *out)
{
    if (!out || !(d1 > 0.0) || !(d2 > 0.0) || !(p >= 0.0 && p <= 1.0)) {
        set_error("F quantile: needs d1, d2 > 0 and p in [0, 1]");
        return SRED_ERR_ARG;
    }
    *out = f_quantile(d1, d2, p);
    return SRED_OK;
}

int sred_simple_regression(const double *x, const double *y, int64_t n, double *slope, double *intercept)
{
    if (n < 0 || (n > 0 && (!x || !y)) || !slope || !intercept) {
        set_error("null buffer or negative size");
        return SRED_ERR_ARG;
    }
    simple_regression_host(x, y, n, slope, intercept);
    return SRED_OK;
}

int sred_quadtree_gat_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, int min_size, float alpha,
                             int max_iter, float lts_fraction, double *g0_edc, int *n_leaves, int cap,
                             int32_t *rects, float *means, float *variances)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_img || !g0_edc) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    return quadtree_gat_dev(ctx, d_img, w, h, min_size, alpha, max_iter, lts_fraction, g0_edc, n_leaves, cap, rects,
                            means, variances);
}

int sred_quadtree_gat_2d(sred_ctx *ctx, const float *img, int w, int h, int min_size, float alpha, int max_iter,
                         float lts_fraction, double *g0_edc, int *n_leaves, int cap, int32_t *rects, float *means,
                         float *variances)
{
    SRED_TRY(check_ctx(ctx));
    if (!img || !g0_edc || w <= 0 || h <= 0) {
        set_error("null buffer or empty image");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, img, (int64_t)w * h));
    return quadtree_gat_dev(ctx, ctx->img.as<float>(), w, h, min_size, alpha, max_iter, lts_fraction, g0_edc,
                            n_leaves, cap, rects, means, variances);
}

int sred_gat_tree_2d(sred_ctx *ctx, float *img, int64_t n, double g0, double edc)
{
    SRED_TRY(check_ctx(ctx));
    if (!img || n < 0) {
        set_error("null buffer or negative size");
        return SRED_ERR_ARG;
    }
    if (n == 0) return SRED_OK;
    SRED_TRY(upload(ctx, ctx->img, img, n));
    SRED_TRY(gat_tree_dev(ctx, ctx->img.as<float>(), n, g0, edc));
    return download(ctx, img, ctx->img, n);
}

int sred_gat_tree_2d_dev(sred_ctx *ctx, float *d_img, int64_t n, double g0, double edc)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_img || n < 0) {
        set_error("null buffer or negative size");
        return SRED_ERR_ARG;
    }
    return gat_tree_dev(ctx, d_img, n, g0, edc);
}

int sred_octree_gat_3d_dev(sred_ctx *ctx, const float *d_vol, int w, int h, int d, int min_size, float alpha,
                           int max_iter, float lts_fraction, double *g0_edc, int *n_leaves, int cap, int32_t *boxes,
                           float *means, float *variances)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_vol || !g0_edc) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    return octree_gat_dev(ctx, d_vol, w, h, d, min_size, alpha, max_iter, lts_fraction, g0_edc, n_leaves, cap, boxes,
                          means, variances);
}

int sred_octree_gat_3d(sred_ctx *ctx, const float *vol, int w, int h, int d, int min_size, float alpha, int max_iter,
                       float lts_fraction, double *g0_edc, int *n_leaves, int cap, int32_t *boxes, float *means,
                       float *variances)
{
    SRED_TRY(check_ctx(ctx));
    if (!vol || !g0_edc || w <= 0 || h <= 0 || d <= 0 || (int64_t)w * h * d >= (1LL << 31)) {
        set_error("null buffer, empty volume or more than 2^31 voxels");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, vol, (int64_t)w * h * d));
    return octree_gat_dev(ctx, ctx->img.as<float>(), w, h, d, min_size, alpha, max_iter, lts_fraction, g0_edc,
                          n_leaves, cap, boxes, means, variances);
}

int sred_gat_tree_3d(sred_ctx *ctx, float *vol, int64_t n, float g0, float edc)
{
    SRED_TRY(check_ctx(ctx));
    if (!vol || n < 0) {
        set_error("null buffer or negative size");
        return SRED_ERR_ARG;
    }
    if (n == 0) return SRED_OK;
    SRED_TRY(upload(ctx, ctx->img, vol, n));
    SRED_TRY(gat_tree3d_dev(ctx, ctx->img.as<float>(), n, g0, edc));
    return download(ctx, vol, ctx->img, n);
}

int sred_gat_tree_3d_dev(sred_ctx *ctx, float *d_vol, int64_t n, float g0, float edc)
{
    SRED_TRY(check_ctx(ctx));
    if (!d_vol || n < 0) {
        set_error("null buffer or negative size");
        return SRED_ERR_ARG;
    }
    return gat_tree3d_dev(ctx, d_vol, n, g0, edc);
}

int sred_global_batch_2d(sred_ctx *ctx, int metric, int n, const sred_global_problem *problems,
                         float relevance_constant, int normalize)
{
    SRED_TRY(check_ctx(ctx));
    if (n < 0 || (n > 0 && !problems)) {
        set_error("invalid problem list");
        return SRED_ERR_ARG;
    }
    if (n == 0) return SRED_OK;
    return global_batch_dev(ctx, metric, n, problems, relevance_constant, normalize);
}

int sred_ctx_tuning_ms(sred_ctx *ctx, int metric, int k_steps, float *ms)
{
    SRED_TRY(check_ctx(ctx));
    if (!ms) {
        set_error("null output");
        return SRED_ERR_ARG;
    }
    for (int i = 0; i < 4; i++) ms[i] = 0.0f;
    auto f = ctx->tune_ms.find(std::make_pair(metric, k_steps));
    if (f != ctx->tune_ms.end())
        for (size_t i = 0; i < f->second.size() && i < 4; i++) ms[i] = f->second[i];
    return SRED_OK;
}

int sred_ctx_tuned_panel(sred_ctx *ctx, int metric, int k_steps, int *panel)
{
    SRED_TRY(check_ctx(ctx));
    if (!panel) {
        set_error("null output");
        return SRED_ERR_ARG;
    }
    auto f = ctx->tuned_panel.find(std::make_pair(metric, k_steps));
    *panel = f == ctx->tuned_panel.end() ? 0 : f->second;
    return SRED_OK;
}

int sred_hu_prepare_template_2d(const float *blk, int bw, int bh, float *tmpl_out)
{
    Geometry g;
    SRED_TRY(make_geometry(bw, bh, bw, bh, &g));
    if (!blk || !tmpl_out) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    return hu_prepare_template_host(blk, bw, bh, g.K, tmpl_out);
}

int sred_global_pearson_adjoint_2d(sred_ctx *ctx, const float *img, int w, int h, int bw, int bh,
                                   const float *dL_dmap, float *dL_dimg)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!img || !dL_dmap || !dL_dimg) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    SRED_TRY(upload(ctx, ctx->img, img, g.n()));
    SRED_TRY(upload(ctx, ctx->dmap, dL_dmap, g.n()));
    SRED_TRY(ctx->grad.ensure(sizeof(float) * g.n()));
    SRED_TRY(global_pearson_adjoint_dev(ctx, g, ctx->img.as<float>(), ctx->dmap.as<float>(), ctx->grad.as<float>()));
    SRED_TRY(download(ctx, dL_dimg, ctx->grad, g.n()));
    return SRED_OK;
}

int sred_global_pearson_adjoint_2d_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh,
                                       const float *d_dmap, float *d_grad, float *d_map)
{
    SRED_TRY(check_ctx(ctx));
    Geometry g;
    SRED_TRY(make_geometry(w, h, bw, bh, &g));
    if (!d_img || !d_dmap || !d_grad) {
        set_error("null buffer");
        return SRED_ERR_ARG;
    }
    SRED_TRY(global_pearson_adjoint_dev(ctx, g, d_img, d_dmap, d_grad));
    if (d_map)
        SRED_HIP_CHECK(hipMemcpyAsync(d_map, ctx->out.ptr, sizeof(float) * g.n(), hipMemcpyDeviceToDevice, ctx->stream));
    return SRED_OK;
}

int sred_ctx_last_skipped(sred_ctx *ctx, int64_t *pairs_skipped)
{
    SRED_TRY(check_ctx(ctx));
    unsigned long long v = 0;
    if (ctx->last_skip) {
        SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
        SRED_HIP_CHECK(hipMemcpy(&v, ctx->last_skip, sizeof v, hipMemcpyDeviceToHost));
    }
    if (pairs_skipped) *pairs_skipped = (int64_t)v;
    return SRED_OK;
}

int sred_global_plan_2d(int n_valid, int block_pixels, int metric, int precision, int symmetric, int chunk_cols,
                        int64_t *out)
{
    if (!out) {
        set_error("null output");
        return SRED_ERR_ARG;
    }
    ContractionPlan pl;
    SRED_TRY(plan_contraction(n_valid, block_pixels, metric, precision, symmetric, chunk_cols, &pl));
    out[0] = pl.Nv_pad;
    out[1] = pl.sym ? pl.P : pl.bm;
    out[2] = pl.cols_per_chunk;
    out[3] = pl.nchunks;
    out[4] = pl.grid;
    out[5] = pl.partial_elems;
    out[6] = pl.pairs_evaluated;
    out[7] = (pl.mfma ? 1 : 0) | (pl.sym ? 2 : 0) | (pl.reflect ? 4 : 0);
    return SRED_OK;
}

int sred_ctx_last_contraction(sred_ctx *ctx, float *ms, int *launches, int64_t *pairs_evaluated)
{
    SRED_TRY(check_ctx(ctx));
    float t = 0.0f;
    if (ctx->summed_ms >= 0.0f) {
        t = ctx->summed_ms;
    } else if (ctx->last_launches > 0 && ctx->ev0) {
        SRED_HIP_CHECK(hipEventSynchronize(ctx->ev1));
        SRED_HIP_CHECK(hipEventElapsedTime(&t, ctx->ev0, ctx->ev1));
    }
    if (ms) *ms = t;
    if (launches) *launches = ctx->last_launches;
    if (pairs_evaluated) *pairs_evaluated = ctx->last_pairs;
    return SRED_OK;
}

}  // extern "C"
