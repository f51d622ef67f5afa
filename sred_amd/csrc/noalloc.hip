// noalloc.hip -- block repetition without device allocation.
//
// sred_block_map_2d_dev_noalloc is sred_block_map_2d_dev for callers that
// cannot allocate device memory (tight serving loops over frames, harnesses
// that forbid hipMalloc).  Every scratch buffer of the block pipeline -- the
// expanded template, the stencil's min/max partials, the normalisation and
// mask-count words, and (relevance constant != 0, or NRMSE) the window
// statistics, the relevance mask and the noise tiles -- is carved out of one
// file-scope device array of kScratchBytes per GPU, which lives until the
// process ends.  Its device address is looked up once per device, with that
// device current, when a context is created on it; host code uses only that
// address (the array's name is a host-side stand-in).  The scratch keeps the
// last call's contents and every call on a GPU shares it: calls that overlap
// on one GPU (two streams, two threads) race.
#include <mutex>

#include "sred_internal.h"

namespace sred {

constexpr size_t kScratchBytes = (size_t)64 << 20;  // 64 MiB per GPU
static __device__ __align__(256) unsigned char scratch[kScratchBytes];

namespace {
std::mutex g_scratch_mu;
std::map<int, unsigned char *> g_scratch_ptr;  // device -> address of that device's copy of `scratch`
}  // namespace

// The current device's copy of `scratch` (the caller has made `device` current).
int noalloc_scratch_init(int device, unsigned char **out)
{
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    auto it = g_scratch_ptr.find(device);
    if (it == g_scratch_ptr.end()) {
        unsigned char *p = nullptr;
        SRED_HIP_CHECK(hipGetSymbolAddress((void **)&p, HIP_SYMBOL(scratch)));
        it = g_scratch_ptr.emplace(device, p).first;
    }
    *out = it->second;
    return SRED_OK;
}

namespace {
// Lends regions of the scratch to the context's buffers for one call: the
// lent buffers are fixed (ensure() never reallocates them); the context's own
// buffers and its template cache come back when the lease ends.
struct ScratchLease {
    sred_ctx *ctx;
    unsigned char *base;
    size_t used = 0;
    DevBuf *slots[8];
    DevBuf saved[8];
    int n = 0;
    std::vector<float> saved_tmpl;
    ScratchLease(sred_ctx *c, unsigned char *b) : ctx(c), base(b)
    {
        saved_tmpl.swap(ctx->host_tmpl);  // the lent template buffer holds no known template
    }
    void lend(DevBuf &buf, size_t bytes)
    {
        saved[n] = buf;
        slots[n++] = &buf;
        buf = DevBuf{};
        buf.ptr = base + used;
        buf.bytes = bytes;
        buf.fixed = true;
        used += (bytes + 255) / 256 * 256;
    }
    ~ScratchLease()
    {
        for (int i = 0; i < n; i++) *slots[i] = saved[i];
        ctx->host_tmpl.swap(saved_tmpl);
    }
};

size_t round256(size_t b) { return (b + 255) / 256 * 256; }
}  // namespace

}  // namespace sred

using namespace sred;

extern "C" int sred_block_map_2d_dev_noalloc(sred_ctx *ctx, int metric, const float *d_img, int w, int h,
                                             const float *tmpl, int bw, int bh, float tmpl_mean, float tmpl_std,
                                             float relevance_constant, int normalize, float *d_out)
{
    if (!ctx) {
        set_error("null context");
        return SRED_ERR_ARG;
    }
    SRED_HIP_CHECK(hipSetDevice(ctx->device));
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
    if (!ctx->noalloc_scratch) {
        set_error("context has no scratch address");
        return SRED_ERR_ARG;
    }
    // what the block pipeline asks of each buffer for this call
    const int64_t n = g.n();
    const bool stats = relevance_constant != 0.0f;
    const bool mask = stats || metric == SRED_NRMSE;
    const size_t b_tmpl = sizeof(float4) * (size_t)(g.K + 3);
    const size_t b_parts = sizeof(unsigned int) * 3 * (size_t)block_workgroups(g);
    const size_t b_map = sizeof(float) * (size_t)n;
    const size_t b_tiles = sizeof(float) * (size_t)((w / 8) * (h / 8)) + 64;  // at most this many noise tiles
    const size_t need = round256(b_tmpl) + round256(kMiscBytes) + round256(b_parts) +
                        (stats ? 2 * round256(b_map) + round256(b_tiles) : 0) + (mask ? round256(b_map) : 0);
    if (need > kScratchBytes) {
        set_error("block map of " + std::to_string(w) + "x" + std::to_string(h) + " pixels needs " +
                  std::to_string(need) + " bytes of scratch; the no-allocation entry point holds " +
                  std::to_string(kScratchBytes) + " per GPU");
        return SRED_ERR_ARG;
    }
    ScratchLease lease(ctx, ctx->noalloc_scratch);
    lease.lend(ctx->tmpl, b_tmpl);
    lease.lend(ctx->misc, kMiscBytes);
    lease.lend(ctx->mmpart, b_parts);
    if (stats) {
        lease.lend(ctx->means, b_map);
        lease.lend(ctx->stds, b_map);
        lease.lend(ctx->tile_var, b_tiles);
    }
    if (mask) lease.lend(ctx->mask, b_map);
    return block_pipeline(ctx, g, metric, d_img, tmpl, tmpl_mean, tmpl_std, relevance_constant, normalize, d_out);
}
