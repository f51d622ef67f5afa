// sred_internal.h -- shared declarations of libsred_hip.so (not installed).
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <string>
#include <utility>
#include <vector>

#include "../../include/sred.h"

namespace sred {

// Order-preserving map of floats to unsigned ints (min/max by integer compare).
__device__ __forceinline__ unsigned int f2ord(float f)
{
    unsigned int u = __float_as_uint(f);
    return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
__device__ __forceinline__ float ord2f(unsigned int u)
{
    return __uint_as_float((u & 0x80000000u) ? (u & 0x7fffffffu) : ~u);
}

// bfloat16 by round-to-nearest-even (finite inputs), and back.
__device__ __forceinline__ unsigned short f2bf(float f)
{
    unsigned int u = __float_as_uint(f);
    u += 0x7fffu + ((u >> 16) & 1u);
    return (unsigned short)(u >> 16);
}
__device__ __forceinline__ float bf2f(unsigned short b) { return __uint_as_float((unsigned int)b << 16); }

// Per-thread error text behind sred_last_error().
void set_error(const std::string &msg);

#define SRED_HIP_CHECK(expr)                                                            \
    do {                                                                                \
        hipError_t e_ = (expr);                                                         \
        if (e_ != hipSuccess) {                                                         \
            ::sred::set_error(std::string(#expr) + ": " + hipGetErrorString(e_));       \
            return SRED_ERR_HIP;                                                        \
        }                                                                               \
    } while (0)

#define SRED_TRY(expr)                   \
    do {                                 \
        int rc_ = (expr);                \
        if (rc_ != SRED_OK) return rc_;  \
    } while (0)

constexpr float kEps = 1e-7f;           // Utils.EPSILON, Utils.java:20
constexpr int kMaxK = 3200;             // footprint size cap (63x63 ellipse has 3109 px)
constexpr int kMaxKS = 8;               // split-BF16 kernels: K <= 32 * kMaxKS (larger blocks use FP32)
constexpr int kMaxKSSym = 12;           // the symmetric split-BF16 kernel (one row group beyond 8 k-steps)
// Context scratch `misc`: [0, 1024) relevance-mask count, [1024, 2048)
// normalisation min/max, [2048, 4096) mask check, [4096, ...) min/max
// partials of up to 1024 workgroups (3 words each).
constexpr size_t kMiscBytes = 4096 + 3 * 4 * 1024;

// Growable device scratch buffer owned by a context.
struct DevBuf {
    void *ptr = nullptr;
    size_t bytes = 0;
    bool fixed = false;  // lent memory (no-allocation calls): ensure() never reallocates, release() never frees
    int ensure(size_t n);
    void release();
    template <class T> T *as() const { return static_cast<T *>(ptr); }
};

// Geometry of one 2D problem.
// Geometry of one problem: a w x h image (d = 1, bd = 1, brd = 0) or a
// w x h x d volume with bw x bh x bd blocks (slices of w*h voxels).
struct Geometry {
    int w = 0, h = 0, bw = 0, bh = 0, brw = 0, brh = 0, K = 0;
    int d = 1, bd = 1, brd = 0;
    int64_t n() const { return (int64_t)w * h * d; }
};

int make_geometry(int w, int h, int bw, int bh, Geometry *g);
int make_geometry3d(int w, int h, int d, int bw, int bh, int bd, Geometry *g);
int footprint3d(int bw, int bh, int bd, std::vector<int> *dx, std::vector<int> *dy, std::vector<int> *dz);
// Row-major footprint offsets (dx, dy), reference predicate.
int footprint(int bw, int bh, std::vector<int> *dx, std::vector<int> *dy);
// The double an OpenCL compiler reads from Java's ""+f.
double java_float_literal(float f);

}  // namespace sred

namespace sred {
// Operand and accumulator vectors of the v_mfma_f32_16x16x32_bf16 kernels.
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float floatx4 __attribute__((ext_vector_type(4)));

// v * s for a power of two s (exact in BF16 while every result stays finite)
__device__ __forceinline__ bf16x8 bf16x8_scaled(bf16x8 v, float s)
{
    bf16x8 r;
#pragma unroll
    for (int e = 0; e < 8; e++) r[e] = (__bf16)((float)v[e] * s);
    return r;
}

// Epilogue constants of the global contractions.
struct EpiConst {
    float eps;  // 1e-7
    float c1;   // SSIM 1e-4
    float c2;   // SSIM 9e-4
    float pk;   // Pearson: eps * log2(e)
};
// One problem of a symmetric split-BF16 contraction launch (a persistent
// launch may carry several): its fragment-major block matrix, per-block
// constants and small-sigma flags, and its partial-sum array [npanels][Nv_pad].
struct SymProblem {
    const uint4 *uhi, *ulo;
    const float4 *aux;
    const int *tsm;
    float *partial;
    int Nv_pad, npanels;
    long long item0;  // index of this problem's first panel pair in the launch's work list
    // Rectangular problem (npr > 0): the panel pairs of npr row panels (blocks
    // [0, npr P) of the arrays above) against the column panels starting at
    // block coff (past the rows); row sums go to partial[c][Nv_pad rows],
    // column sums to pcol[p][ldc columns].  npr = 0: the triangle of
    // npanels panels, column sums in partial as well.
    float *pcol = nullptr;
    int ldc = 0, coff = 0, npr = 0;
};
}  // namespace sred

namespace sred {
// Per-image scratch of a batched global call (owned by the context).
struct BatchStore {
    DevBuf means, stds, mask, S, W, counts, valid_idx, aux, tflags, u32, uhi, ulo, offs2, wmom, wpart, sortbuf;
    DevBuf tile_var, count;  // noise tiles, relevance-mask pixel count
    int offs2_key[3] = {-1, -1, -1};
    ~BatchStore();
    void release();
    struct GlobalBufs bufs();
};
}  // namespace sred

struct sred_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    int chunk_cols = 0;
    int symmetric = 1;  // MFMA contraction over upper-triangle panel pairs
    // device scratch
    sred::DevBuf img, means, stds, mask, S, W, out, dmap, grad;          // pixel maps (host-pointer paths)
    sred::DevBuf tile_var, counts, valid_idx, aux, tflags, u32, uhi, ulo, partial, misc, offs, offs2, tmpl;
    sred::DevBuf wmom, wpart;  // weight-sum moments
    sred::DevBuf sortbuf;      // valid blocks sorted by sigma
    sred::DevBuf mmpart;       // per-workgroup min/max partials of a block map
    sred::DevBuf adj_v;        // transposed split-BF16 fragments of the adjoint
    sred::DevBuf adj_f;        // the adjoint's sigma-only terms (FP64 per block) and max |b|
    sred::DevBuf pm_rows, pm_cols, pm_part, pm_spec;  // all-pairs vector maps (Hu, phase correlation)
    sred::DevBuf qt_res, qt_nodes, qt_leaves, qt_seg, qt_tmp;  // quadtree GAT estimate
    sred::DevBuf stack_work, seg_mm;                           // time-lapse stacks
    sred::DevBuf tmpl3d, offs3d;  // reference block of a volume; ellipsoid offsets dz*w*h + dy*w + dx
    int offs3d_key[5] = {-1, -1, -1, -1, -1};  // (bw, bh, bd, w, h) of the offsets in offs3d
    std::vector<float> host_tmpl3d;            // reference block last copied into tmpl3d
    sred::DevBuf symtab;       // problem table of the symmetric contraction
    std::vector<sred::SymProblem> host_symtab;
    sred::DevBuf acc64;        // FP64 row totals of a contraction run in super-panels
    int64_t partial_mb = 16384;  // partial-sum memory of one global call before it runs in super-panels
    // batched global calls
    std::vector<sred::BatchStore *> batch_store;
    sred::DevBuf batch_part, batch_tab;  // partial sums of a wave; problem tables + item counter
    size_t batch_ctr_off = 0;
    std::map<std::pair<int, int>, int> tuned_panel;  // (metric, KS) -> panel size
    std::map<std::pair<int, int>, std::vector<float>> tune_ms;  // (metric, KS) -> ms per candidate panel (0: not timed)
    int autotune = 1;
    int batch_panel = 0;       // forced panel size (0: autotuner)
    int64_t batch_mb = 16384;  // partial-sum memory per wave
    int fit_timing = 0;        // GAT fit: HIP events around every objective launch
    int offs_key[3] = {-1, -1, -1}, offs2_key[3] = {-1, -1, -1};
    std::vector<float> host_tile_var;
    std::vector<float> host_tmpl;  // expanded template last copied into `tmpl` (skips repeated uploads)
    // measurement of the last global contraction
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    float last_ms = 0.0f;
    float summed_ms = -1.0f;  // >= 0: the dominant kernel's time summed over the launches of the last call
    int last_launches = 0;
    int64_t last_pairs = 0;
    int last_nv = 0;  // valid blocks of the last global accumulation
    float last_sigma_max = 0.0f;  // largest valid block sigma of the last global accumulation
    unsigned long long *last_skip = nullptr;  // device counter of pairs skipped by the last symmetric contraction
    double *h_fit = nullptr;  // pinned host doubles the GAT objective kernel writes its values to
    unsigned char *noalloc_scratch = nullptr;  // this device's copy of the no-allocation scratch array
    void *h_stage = nullptr;  // pinned host staging of batched readbacks
    size_t h_stage_bytes = 0;
};

namespace sred {

// Bracket the dominant kernel of a call with the context's HIP events on its
// stream (read back by sred_ctx_last_contraction); units = work items of it.
int time_begin(sred_ctx *ctx);
int time_end(sred_ctx *ctx, int64_t units);
// Upload the footprint offsets dy*pitch+dx (cached per (bw, bh, pitch)).
int upload_offsets(sred_ctx *ctx, const Geometry &g, int pitch, DevBuf *buf, int *key);
// Zero device memory with a kernel (graph-capture safe, unlike memset nodes).
int zero_dev(sred_ctx *ctx, void *ptr, size_t bytes);

// ---- launchers (stream = ctx->stream) ----
// Local statistics, reference float order (kernelGetLocalStatistics2D.cl).
int launch_local_stats(sred_ctx *ctx, const Geometry &g, const float *d_img, float *d_means, float *d_stds,
                       bool var_out = false);
// Java float noise variance (Utils.getMeanNoiseVar2D): tile variances on the
// GPU, sort and mean of the lowest 3% on the host (synchronises).
int noise_var_dev(sred_ctx *ctx, const float *d_img, int w, int h, float *noise_var, bool double_sums = false);
// Relevance mask + Java float nPixels (device); n_pixels returned on host
// (synchronises).
int relevance_mask_dev(sred_ctx *ctx, const Geometry &g, const float *d_stds, float threshold, float *d_mask,
                       float *n_pixels);

struct ContractionPlan {
    bool mfma, sym;
    // split-BF16 blocks of K = 32m + 1 pixels are reflected into K - 1
    // dimensions (the centred vectors are orthogonal to the all-ones vector),
    // so they take m k-steps instead of m + 1
    bool reflect;
    int KS, K4, P, Nv_pad, bm, nrowtiles, cols_per_chunk, nchunks;
    int64_t grid, partial_elems, pairs_evaluated;
};
int plan_contraction(int Nv, int K, int metric, int precision, int symmetric, int chunk_cols, ContractionPlan *pl);
// Same with the symmetric panel size forced (panel > 0: multiple of 256, <= 2048).
int plan_contraction_p(int Nv, int K, int metric, int precision, int symmetric, int chunk_cols, int panel,
                       ContractionPlan *pl);

struct GlobalParams {
    int metric = SRED_PEARSON;
    int precision = SRED_FP32;
    float filter_param = 0.0f;  // noise variance
    float threshold = 0.0f;     // noise variance * relevance constant
    bool skip_flat_rows = false;  // flat reference blocks score 0 (the 3D kernels skip them)
    bool nrmse_weighted = false;  // NRMSE adds sqrt(mse) * weight (kernelGetGlobalNrmse3D.cl:111)
    bool drop_flat = false;       // flat blocks are neither rows nor columns (kernelGetGlobalAbsDiffStds3D.cl:36,53)
};
// Contraction template id of NRMSE with weighted terms (next to the public
// metric ids 0-3).
constexpr int kNrmseWeighted = 4;
// Scratch of one global problem (pointers to buffers owned by a context).
struct GlobalBufs {
    DevBuf *counts, *valid_idx, *aux, *tflags, *u32, *uhi, *ulo, *offs2;
    int *offs2_key;
    DevBuf *wmom, *wpart, *sortbuf;
};
GlobalBufs ctx_global_bufs(sred_ctx *ctx);
// A prepared global problem (see global_prepare).
struct GlobalPrepared {
    int Nv = 0, Nv_rows = 0;
    float flat_count = 0.0f;
    bool zero_flat_rows = false;  // every metric's flat rows take flat_count (0)
    double tscale_d = 0.0;
    ContractionPlan pl{};
    const int *d_idx = nullptr;
    const float4 *d_aux = nullptr;
    const int *d_tsm = nullptr;
    const uint4 *uhi = nullptr, *ulo = nullptr;
    const float *u32 = nullptr;
};
int global_prepare(sred_ctx *ctx, const GlobalBufs &b, const Geometry &g, const GlobalParams &p, const float *d_img,
                   const float *d_means, const float *d_stds, float *d_S, float *d_W, int symmetric, int panel,
                   int pad_rows, GlobalPrepared *out);
// global_prepare in two parts around its one synchronisation: the
// compaction (enqueued, totals copied to h_tot[2]) and the rest once h_tot
// holds them (batched callers synchronise once for many problems).
int global_prepare_front(sred_ctx *ctx, const GlobalBufs &b, const Geometry &g, const GlobalParams &p,
                         const float *d_stds, int2 *h_tot);
int global_prepare_back(sred_ctx *ctx, const GlobalBufs &b, const Geometry &g, const GlobalParams &p,
                        const float *d_img, const float *d_means, const float *d_stds, float *d_S, float *d_W,
                        int symmetric, int panel, int pad_rows, const int2 *tot, GlobalPrepared *out);
// The noise variance in two parts: the tile variances enqueued (into `tiles`,
// copied with the non-finite flag to h_tiles[0..nb], host memory), then the
// host sort and mean once the stream has been synchronised.  noise_tiles
// gives nb (or -1 for an image smaller than one tile).
int noise_tiles(int w, int h, int *tbw, int *tbh);
int noise_var_enqueue(sred_ctx *ctx, DevBuf &tiles, const float *d_img, int w, int h, float *h_tiles,
                      bool double_sums = false);
int noise_var_finish(float *h_tiles, int nb, float *noise_var);
// The relevance mask with its pixel count copied to *h_count (enqueued);
// n_pixels_of turns the count into Java's float nPixels.
int relevance_mask_enqueue(sred_ctx *ctx, const Geometry &g, const float *d_stds, float threshold, float *d_mask,
                           unsigned long long *d_count, unsigned long long *h_count);
float n_pixels_of(unsigned long long count);
int global_reduce_dev(sred_ctx *ctx, const GlobalPrepared &r, int metric, const float *d_part, int nchunks,
                      int Nv_pad, const float *d_stds, float *d_S);
EpiConst epi_const();
// Accumulators S, Wsum (float maps, w*h) of the global kernels.
int global_accum_dev(sred_ctx *ctx, const Geometry &g, const GlobalParams &p, const float *d_img,
                     const float *d_means, const float *d_stds, float *d_S, float *d_W);
// The valid-block list d_idx[Nv] reordered by ascending sigma (stable), in
// buf; *d_sorted points into buf.
int sort_valid_by_sigma(sred_ctx *ctx, DevBuf &buf, const float *d_stds, const int *d_idx, int Nv, int **d_sorted);
// Weight sums W (float map entries idx[v], v < Nv) of the global kernels:
// sum over valid j of exp2(-(t_i - t_j)^2), t = sigma * tscale, by a binned
// Hermite expansion in FP64; sigma_max bounds the valid blocks' sigmas.
int weight_sums_dev(sred_ctx *ctx, DevBuf &wmom, DevBuf &wpart, const float *d_stds, const int *d_idx, int Nv,
                    double tscale, float sigma_max, float *d_W);
// The sigma-only term of the global Pearson adjoint, F_v = sum_j (b_v + b_j)
// w_vj (t_v - t_j) over the Nv valid blocks in adjoint order (info2[v].z = b_v,
// *d_bmax = bits of max |b|), by the same expansion; FP64 results.
int adjoint_weight_terms_dev(sred_ctx *ctx, DevBuf &wmom, DevBuf &wpart, const float *d_stds, const int *d_idx,
                             const float4 *d_info2, const unsigned int *d_bmax, int Nv, double tscale,
                             float sigma_max, double *d_F);
// Persistent symmetric split-BF16 contraction over nitems panel pairs of the
// nprob problems in d_probs (device table), all with K_pad = 32*KS and panel
// size P (metric Pearson, SSIM or NRMSE).
// d_next: a 64-bit device counter the launch zeroes and its workgroups claim items from.
// rect: the table may hold rectangular problems (SymProblem::npr > 0).
int launch_sym_metric(int metric, int KS, long long nitems, hipStream_t st, const SymProblem *d_probs, int nprob,
                      unsigned long long *d_next, int P, EpiConst ec, bool rect = false);
// Split-BF16 gather of the valid blocks, fragment-major (Uhi/Ulo[Nv_pad/16][KS][64]).
void launch_gather_bf16(sred_ctx *ctx, int metric, const float *d_img, const float *d_means, const float *d_stds,
                        const int *d_idx, int Nv, int Nv_pad, int K, int KS, const int *d_offs, float inv_sqrt_km1,
                        float inv_sqrt_k, uint4 *Uhi, uint4 *Ulo);
// FP32 gather of the valid blocks, k-major (UT[k][v]).
void launch_gather_f32(sred_ctx *ctx, int metric, const float *d_img, const float *d_means, const float *d_stds,
                       const int *d_idx, int Nv, int Nv_pad, int K, int K4, const int *d_offs, float inv_sqrt_km1,
                       float inv_sqrt_k, float *UT);
// Gradient of sum(dmap * map) of the global Pearson map (relevance 0,
// normalised) with respect to the pixels (device buffers).
int global_pearson_adjoint_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, const float *d_dmap,
                               float *d_grad);
// Host-finalize of CLUtils.getGlobal*2D on device maps.
int global_finalize_dev(sred_ctx *ctx, const Geometry &g, int metric, const float *d_S, const float *d_W,
                        const float *d_mask, float n_pixels, float c, int normalize, float *d_out);
// Block-repetition raw kernel output (before host post-processing); window
// means/stds are written when both output pointers are non-NULL.
// Normalisation without mask of a map whose interior min/max partials (3
// words each: min, max, 0 as ordered ints) are given; init_max_min_value as
// normalize_map_dev.
int normalize_frames_from_parts_dev(sred_ctx *ctx, const Geometry &g, float *d_maps, int nframes,
                                    const unsigned int *d_parts, int nparts, int init_max_min_value);
int normalize_map_from_parts_dev(sred_ctx *ctx, const Geometry &g, float *d_map, const unsigned int *d_parts,
                                 int nparts, int init_max_min_value);
// d_mmpart (optional): per-workgroup interior min/max partials of the output
// (block_workgroups(g) of them), *nparts receives their count.
int64_t block_workgroups(const Geometry &g);
int block_raw_dev(sred_ctx *ctx, const Geometry &g, int metric, const float *d_img, float *d_means_out,
                  float *d_stds_out, const float *h_tmpl, float tmpl_mean, float tmpl_std, float *d_out,
                  unsigned int *d_mmpart = nullptr, int *nparts = nullptr, int nframes = 1);
// Batched global maps (see batch.hip).
int global_batch_dev(sred_ctx *ctx, int metric, int n, const sred_global_problem *probs, float c, int normalize);
// Hu template map (kernelGetPatchHu2D.cl) of a device image; h_tmpl is the
// host template rectangle.  Synchronises.
int block_hu_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, const float *h_tmpl, float *d_out);
// hu1 of a template rectangle (kernelGetPatchHu2D.cl:36-101), host.
double hu1_template_host(const float *tmpl, int bw, int bh, double eps);
// Hu template preparation (the window steps of kernelGetPatchHu2D.cl:127-182), host.
int hu_prepare_template_host(const float *blk, int bw, int bh, int K, float *out);
// Per-pixel Hu invariant nu20 + nu02 (kernelGetHuMap.cl:146-172), border 0.
int hu_invariants_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, const float *d_means, float *d_out);
// Global Hu map (kernelGetHuMap.cl) and phase correlation map (textbook
// specification of kernelGetPhaseCorrelationMap.cl) of a device image with
// its local statistics; rows: optional host list of reference pixels.
// Synchronise.
int global_hu_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, const float *d_means, const float *d_stds,
                  const int32_t *rows, int64_t nrows, float *d_out);
int phase_map_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, const float *d_means, const float *d_stds,
                  const int32_t *rows, int64_t nrows, float *d_out);
// Variance stabilisation (vst.hip): getGAT in place on a device image; the
// GATMinimizer2D objective of one parameter set; the simplex fit.
int gat_dev(sred_ctx *ctx, float *d_img, int64_t n, double gain, double sigma, double offset);
int gat_objective_dev(sred_ctx *ctx, const float *d_img, int w, int h, double gain, double sigma, double offset,
                      double *error);
int gat_minimize_dev(sred_ctx *ctx, const float *d_img, int w, int h, double gain0, double sigma0, double offset0,
                     int max_iter, double *params, double *error);
// GATMinimizer3D's objective and fit on a w x h x d volume.
int gat_objective3d_dev(sred_ctx *ctx, const float *d_vol, int w, int h, int d, double gain, double sigma,
                        double offset, double *error);
int gat_minimize3d_dev(sred_ctx *ctx, const float *d_vol, int w, int h, int d, double gain0, double sigma0,
                       double offset0, int max_iter, double *params, double *error);
// The GAT kernel with ref given directly: v <- (g v + ref < 0) ? 0 : (2/g) sqrt(g v + ref).
int gat_ref_dev(sred_ctx *ctx, float *d_img, int64_t n, double gain, double ref);
// Quadtree GAT estimate (quadtree.hip): leaves, robust means, LTS variances
// and (g0, eDC); applyGATtree; the host F quantile and regression it uses.
int quadtree_gat_dev(sred_ctx *ctx, const float *d_img, int w, int h, int min_size, float alpha, int max_iter,
                     float frac, double *g0_edc, int *n_leaves, int cap, int32_t *rects, float *means, float *vars);
int gat_tree_dev(sred_ctx *ctx, float *d_img, int64_t n, double g0, double edc);
// OcTree_'s estimate of a w x h x d volume (boxes: 6 ints per leaf) and its transform.
int octree_gat_dev(sred_ctx *ctx, const float *d_vol, int w, int h, int d, int min_size, float alpha, int max_iter,
                   float frac, double *g0_edc, int *n_leaves, int cap, int32_t *boxes, float *means, float *vars);
int gat_tree3d_dev(sred_ctx *ctx, float *d_vol, int64_t n, float g0, float edc);
double f_quantile(double d1, double d2, double p);
// nseg min-max normalisations (v - min) / (max - min + EPSILON) of
// consecutive segments of seg floats (stack.hip); d_in may equal d_out.
int seg_normalize_dev(sred_ctx *ctx, const float *d_in, float *d_out, int64_t seg, int nseg);
// Exact (min, max) of nseg segments as order-preserving keys, mm[2*s], mm[2*s+1].
int seg_minmax_dev(sred_ctx *ctx, const float *d_in, int64_t seg, int nseg, unsigned int *mm);
// Volumes (vol.hip): 3D local statistics, the 3D relevance mask, and the
// global map of a volume Geometry.
int local_stats3d_dev(sred_ctx *ctx, const Geometry &g, const float *d_img, float *d_means, float *d_stds);
int vol_relevance_mask_dev(sred_ctx *ctx, const Geometry &g, const float *d_stds, float c, float *d_mask,
                           float *noise_var, float *threshold, float *n_pixels);
int global_map3d_dev(sred_ctx *ctx, const Geometry &g, int metric, int precision, const float *d_img, float c,
                     int normalize, float *d_out, float *d_means, float *d_stds, float *d_mask, float *d_S,
                     float *d_W);
// BlockRepetition3D_'s map of one reference block (h_tmpl: K host floats of
// Utils.getReferenceBlock3D) over a volume; d_means/d_stds/d_mask receive
// the statistics and (when c > 0) the 3D mask.
// Uploads the ellipsoid offsets of g into ctx->offs3d unless they are there.
int upload_offsets3d(sred_ctx *ctx, const Geometry &g);
int block_map3d_dev(sred_ctx *ctx, const Geometry &g, int metric, const float *d_img, const float *h_tmpl,
                    float tmpl_mean, float tmpl_std, float c, int normalize, float *d_out, float *d_means,
                    float *d_stds, float *d_mask);
// Relevance-mask tools (tools.hip).
int tool_prep_dev(sred_ctx *ctx, float *d_img, int w, int h, int max_iter, double *params);
int relevance_mask_tool_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh, float c, float *d_mask,
                            float *noise_var, float *threshold);
int optimise_relevance_mask_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh, float step,
                                int64_t max_steps, float *all_ones, float *all_zeros);
int relevance_mask_stack_dev(sred_ctx *ctx, const float *d_img, int w, int h, int bw, int bh, float lower,
                             float upper, float step, int cap, float *d_masks, float *constants, int *n_masks);
// The stack normalised as a whole, then each of its nseg frames of seg floats.
int stack_frames_normalize_dev(sred_ctx *ctx, const float *d_in, float *d_out, int64_t seg, int nseg);
void simple_regression_host(const double *x, const double *y, int64_t n, double *slope, double *intercept);
// The block pipeline of sred_block_map_2d_dev (statistics, template
// comparison, mask and the host steps of CLUtils.getBlock*2D) on device buffers.
int block_pipeline(sred_ctx *ctx, const Geometry &g, int metric, const float *d_img, const float *tmpl,
                   float tmpl_mean, float tmpl_std, float c, int normalize, float *d_out);
// Address of `device`'s copy of the no-allocation scratch array (looked up
// once per device, with the device current).
int noalloc_scratch_init(int device, unsigned char **out);
// Masked min/max normalisation (Utils.normalizeImage2D) on a device map.
// mask may be NULL.  init_max selects -FLT_MAX (0) or Float.MIN_VALUE (1).
int normalize_map_dev(sred_ctx *ctx, const Geometry &g, float *d_map, const float *d_mask, int init_max_min_value);
int apply_mask_dev(sred_ctx *ctx, const Geometry &g, float *d_map, const float *d_mask);
int interior_ones_mask_dev(sred_ctx *ctx, const Geometry &g, float *d_mask);
// In-place inversion where mask > 0: rep==0 -> 1 else 1/rep (CLUtils.java:1997-2010).
int invert_masked_dev(sred_ctx *ctx, const Geometry &g, float *d_map, const float *d_mask);
// 1 - x on the interior (CLUtils.java:962-966).
int one_minus_dev(sred_ctx *ctx, const Geometry &g, float *d_map);
// Check that a device mask is binary on the interior.
int mask_is_binary_dev(sred_ctx *ctx, const Geometry &g, const float *d_mask, int *is_binary);

}  // namespace sred
