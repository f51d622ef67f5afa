// batch.hip -- batched global repetition: many images of uneven sizes at
// several block sizes in one HIP graph launch.
//
// No counterpart in the reference (SReD runs one image per command; stacks
// loop in Java, BlockRepetition2D_.java:95-137).  Per image the pipeline is
// that of sred_global_map_2d_dev (GlobalRepetition2D_.java:124-145) with the
// symmetric split-BF16 contraction:
//
//   front   (three stages over all images, two synchronisations per batch):
//           statistics and noise tiles; after the first synchronisation the
//           noise variances (host sort), relevance masks and compactions;
//           after the second nPixels, weight sums, per-block scalars and the
//           gathered block matrices (padded to a multiple of 2048 rows so any
//           panel size up to 2048 applies);
//   graph   (one hipGraphLaunch): per block-size class (K_pad = 32*KS) and
//           memory wave, ONE persistent contraction kernel whose workgroups
//           walk the panel pairs of every image of the class (SymProblem
//           table); then each image's FP64 reduction, finalize, mask and
//           normalisation.
//
// Panel size per (metric, KS): SRED_OPT_BATCH_PANEL if set, else the
// autotuner's choice (cached in the context): on the first batch holding a
// class it times one persistent launch per candidate (256..2048, limited by
// the partial-sum memory of the largest image) over a subset of the class's
// images and keeps the fastest.  Waves bound the partial-sum memory
// (npanels x Nv_pad floats per image) by SRED_OPT_BATCH_MEMORY_MB; the waves
// run one after the other inside the graph and share one buffer.
#include <algorithm>
#include <cmath>
#include <map>
#include <vector>

#include "sred_internal.h"

namespace sred {

constexpr int kPadRows = 2048;
constexpr int kPanels[] = {256, 512, 1024, 2048};

BatchStore::~BatchStore() { release(); }
void BatchStore::release()
{
    DevBuf *b[] = {&means, &stds, &mask, &S, &W, &counts, &valid_idx, &aux, &tflags, &u32, &uhi, &ulo, &offs2,
                   &wmom, &wpart, &sortbuf, &tile_var, &count};
    for (DevBuf *x : b) x->release();
    offs2_key[0] = offs2_key[1] = offs2_key[2] = -1;
}
GlobalBufs BatchStore::bufs()
{
    return GlobalBufs{&counts, &valid_idx, &aux, &tflags, &u32, &uhi, &ulo, &offs2, offs2_key, &wmom, &wpart,
                      &sortbuf};
}

namespace {

struct Item {
    Geometry g;
    const float *d_img = nullptr;
    float *d_out = nullptr;
    BatchStore *st = nullptr;
    GlobalPrepared r;
    float n_pixels = 0.0f;
    // with the class panel size
    int Nv_pad = 0, npanels = 0;
    int64_t part_off = 0;  // floats into the wave's partial buffer
};

int64_t partial_floats(int Nv, int P)
{
    const int64_t Nv_pad = (int64_t)(Nv + P - 1) / P * P;
    return Nv_pad / P * Nv_pad;
}

struct Launch {
    int KS, P, nprob;
    int64_t tab_off, nitems;
    std::vector<int> items;  // indices into the item list
};

// SymProblem entries of `ids` with panel P, partial offsets from part_off
// (updated), appended to tab.
int64_t fill_table(std::vector<Item> &it, const std::vector<int> &ids, int P, float *d_part,
                   std::vector<SymProblem> &tab)
{
    int64_t item0 = 0, off = 0;
    for (int i : ids) {
        Item &x = it[i];
        x.Nv_pad = (x.r.Nv + P - 1) / P * P;
        x.npanels = x.Nv_pad / P;
        x.part_off = off;
        tab.push_back(SymProblem{x.r.uhi, x.r.ulo, x.r.d_aux, x.r.d_tsm, d_part + off, x.Nv_pad, x.npanels, item0});
        item0 += (int64_t)x.npanels * (x.npanels + 1) / 2;
        off += (int64_t)x.npanels * x.Nv_pad;
    }
    return item0;
}

// The table in ctx->batch_tab, followed by the launches' item counter.
size_t counter_offset(size_t entries) { return (sizeof(SymProblem) * entries + 255) / 256 * 256; }
unsigned long long *batch_counter(sred_ctx *ctx)
{
    return reinterpret_cast<unsigned long long *>(ctx->batch_tab.as<char>() + ctx->batch_ctr_off);
}
int upload_table(sred_ctx *ctx, const std::vector<SymProblem> &tab)
{
    ctx->batch_ctr_off = counter_offset(tab.size());
    SRED_TRY(ctx->batch_tab.ensure(ctx->batch_ctr_off + 256));
    if (!tab.empty())
        SRED_HIP_CHECK(hipMemcpy(ctx->batch_tab.ptr, tab.data(), sizeof(SymProblem) * tab.size(),
                                 hipMemcpyHostToDevice));
    return SRED_OK;
}

// Panel size of a class (see the file comment).
int class_panel(sred_ctx *ctx, int metric, int KS, std::vector<Item> &it, const std::vector<int> &ids, int *P_out)
{
    if (ctx->batch_panel > 0) {
        *P_out = ctx->batch_panel;
        return SRED_OK;
    }
    const auto key = std::make_pair(metric, KS);
    auto f = ctx->tuned_panel.find(key);
    if (f != ctx->tuned_panel.end()) {
        *P_out = f->second;
        return SRED_OK;
    }
    int max_nv = 0;
    for (int i : ids) max_nv = std::max(max_nv, it[i].r.Nv);
    const int64_t budget = (int64_t)ctx->batch_mb << 18;  // floats
    std::vector<int> cand;
    for (int P : kPanels)
        if (partial_floats(max_nv, P) <= budget) cand.push_back(P);
    if (cand.empty()) cand.push_back(2048);
    if (!ctx->autotune || cand.size() == 1) {
        *P_out = cand.back();
        ctx->tuned_panel[key] = *P_out;
        return SRED_OK;
    }
    // tuning subset: images in order until ~1e11 pairs, within the budget at the smallest candidate
    std::vector<int> sub;
    double pairs = 0.0;
    int64_t mem = 0;
    for (int i : ids) {
        const int64_t m = partial_floats(it[i].r.Nv, cand.front());
        if (!sub.empty() && (pairs >= 1e11 || mem + m > budget)) break;
        sub.push_back(i);
        pairs += (double)it[i].r.Nv * it[i].r.Nv;
        mem += m;
    }
    SRED_TRY(ctx->batch_part.ensure(sizeof(float) * (size_t)std::max<int64_t>(mem, 1)));
    const EpiConst ec = epi_const();
    float best_ms = 1e30f;
    int best = cand.back();
    std::vector<float> &times = ctx->tune_ms[key];
    times.assign(sizeof kPanels / sizeof kPanels[0], 0.0f);
    for (int P : cand) {
        std::vector<SymProblem> tab;
        const int64_t nitems = fill_table(it, sub, P, ctx->batch_part.as<float>(), tab);
        SRED_TRY(upload_table(ctx, tab));
        SRED_TRY(time_begin(ctx));
        SRED_TRY(launch_sym_metric(metric, KS, nitems, ctx->stream, ctx->batch_tab.as<SymProblem>(), (int)tab.size(),
                                   batch_counter(ctx), P, ec));
        SRED_TRY(time_end(ctx, nitems));
        float ms = 0.0f;
        SRED_HIP_CHECK(hipEventSynchronize(ctx->ev1));
        SRED_HIP_CHECK(hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        for (size_t c = 0; c < times.size(); c++)
            if (kPanels[c] == P) times[c] = ms;
        if (ms < best_ms) {
            best_ms = ms;
            best = P;
        }
    }
    ctx->tuned_panel[key] = best;
    *P_out = best;
    return SRED_OK;
}

}  // namespace

int global_batch_dev(sred_ctx *ctx, int metric, int n, const sred_global_problem *probs, float c, int normalize)
{
    if (metric < 0 || metric > 2) {
        set_error("the batched global map takes Pearson, SSIM or NRMSE");
        return metric == SRED_ABSDIFF ? SRED_ERR_ARG : SRED_ERR_UNKNOWN_METRIC;
    }
    std::vector<Item> it(n);
    ctx->last_skip = nullptr;
    while ((int)ctx->batch_store.size() < n) ctx->batch_store.push_back(new BatchStore());
    // ---- front: three stages over all images, two host synchronisations
    std::vector<int> nb(n);
    std::vector<size_t> tile_off(n);
    size_t tile_floats = 0;
    for (int i = 0; i < n; i++) {
        Item &x = it[i];
        SRED_TRY(make_geometry(probs[i].w, probs[i].h, probs[i].bw, probs[i].bh, &x.g));
        if (!probs[i].d_img || !probs[i].d_out) {
            set_error("null buffer in batch problem " + std::to_string(i));
            return SRED_ERR_ARG;
        }
        // K = 32m + 1 is stored reflected in K - 1 dimensions (plan_contraction_p)
        const int kd = (x.g.K > 1 && (x.g.K - 1) % 32 == 0) ? x.g.K - 1 : x.g.K;
        if ((kd + 31) / 32 > kMaxKSSym) {
            set_error("the batched split-BF16 contraction supports block sizes up to K = 385 pixels");
            return SRED_ERR_ARG;
        }
        int tbw, tbh;
        nb[i] = noise_tiles(x.g.w, x.g.h, &tbw, &tbh);
        if (nb[i] < 0) return SRED_ERR_ARG;
        tile_off[i] = tile_floats;
        tile_floats += (size_t)(nb[i] + 1 + 3) / 4 * 4;
        x.d_img = probs[i].d_img;
        x.d_out = probs[i].d_out;
        x.st = ctx->batch_store[i];
    }
    // pinned host staging: tile variances, mask counts, compaction totals
    const size_t stage_bytes = sizeof(float) * tile_floats + sizeof(unsigned long long) * n + 2 * sizeof(int2) * n;
    if (ctx->h_stage_bytes < stage_bytes) {
        if (ctx->h_stage) SRED_HIP_CHECK(hipHostFree(ctx->h_stage));
        ctx->h_stage = nullptr;
        ctx->h_stage_bytes = 0;
        SRED_HIP_CHECK(hipHostMalloc(&ctx->h_stage, stage_bytes, hipHostMallocDefault));
        ctx->h_stage_bytes = stage_bytes;
    }
    float *h_tiles = static_cast<float *>(ctx->h_stage);
    unsigned long long *h_count = reinterpret_cast<unsigned long long *>(h_tiles + tile_floats);
    int2 *h_tot = reinterpret_cast<int2 *>(h_count + n);
    // stage 1: local statistics and noise tiles of every image
    for (int i = 0; i < n; i++) {
        Item &x = it[i];
        const int64_t np = x.g.n();
        BatchStore &st = *x.st;
        SRED_TRY(st.means.ensure(sizeof(float) * np));
        SRED_TRY(st.stds.ensure(sizeof(float) * np));
        SRED_TRY(st.mask.ensure(sizeof(float) * np));
        SRED_TRY(st.S.ensure(sizeof(float) * np));
        SRED_TRY(st.W.ensure(sizeof(float) * np));
        SRED_TRY(st.count.ensure(sizeof(unsigned long long)));
        SRED_TRY(launch_local_stats(ctx, x.g, x.d_img, st.means.as<float>(), st.stds.as<float>()));
        SRED_TRY(noise_var_enqueue(ctx, st.tile_var, x.d_img, x.g.w, x.g.h, h_tiles + tile_off[i]));
    }
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    // stage 2: noise variances (host), relevance masks with their counts, compactions
    std::vector<GlobalParams> par(n);
    for (int i = 0; i < n; i++) {
        Item &x = it[i];
        BatchStore &st = *x.st;
        float nv;
        SRED_TRY(noise_var_finish(h_tiles + tile_off[i], nb[i], &nv));
        const float thr = nv * c;
        SRED_TRY(relevance_mask_enqueue(ctx, x.g, st.stds.as<float>(), thr, st.mask.as<float>(),
                                        st.count.as<unsigned long long>(), h_count + i));
        GlobalParams &p = par[i];
        p.metric = metric;
        p.precision = SRED_BF16X3;
        p.filter_param = nv;
        p.threshold = thr;
        SRED_TRY(global_prepare_front(ctx, st.bufs(), x.g, p, st.stds.as<float>(), h_tot + 2 * i));
    }
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));
    // stage 3: nPixels, weight sums, per-block scalars, gathered block matrices
    for (int i = 0; i < n; i++) {
        Item &x = it[i];
        BatchStore &st = *x.st;
        x.n_pixels = n_pixels_of(h_count[i]);
        SRED_TRY(global_prepare_back(ctx, st.bufs(), x.g, par[i], x.d_img, st.means.as<float>(), st.stds.as<float>(),
                                     st.S.as<float>(), st.W.as<float>(), 1, 0, kPadRows, h_tot + 2 * i, &x.r));
    }
    // ---- classes, panel sizes, waves
    std::map<int, std::vector<int>> cls;  // KS -> images with Nv > 0
    for (int i = 0; i < n; i++)
        if (it[i].r.Nv > 0) cls[it[i].r.pl.KS].push_back(i);
    const int64_t budget = (int64_t)ctx->batch_mb << 18;
    std::vector<Launch> launches;
    int64_t max_wave = 1;
    for (auto &kv : cls) {
        int P;
        SRED_TRY(class_panel(ctx, metric, kv.first, it, kv.second, &P));
        std::vector<int> wave;
        int64_t mem = 0;
        auto flush = [&]() {
            if (wave.empty()) return;
            launches.push_back(Launch{kv.first, P, (int)wave.size(), 0, 0, wave});
            max_wave = std::max(max_wave, mem);
            wave.clear();
            mem = 0;
        };
        for (int i : kv.second) {
            const int64_t m = partial_floats(it[i].r.Nv, P);
            if (!wave.empty() && mem + m > budget) flush();
            wave.push_back(i);
            mem += m;
        }
        flush();
    }
    // every wave's partial sums share one buffer; the tables point into it
    SRED_TRY(ctx->batch_part.ensure(sizeof(float) * (size_t)max_wave));
    float *d_part = ctx->batch_part.as<float>();
    std::vector<SymProblem> tab;
    for (Launch &L : launches) {
        L.tab_off = (int64_t)tab.size();
        L.nitems = fill_table(it, L.items, L.P, d_part, tab);
    }
    SRED_TRY(upload_table(ctx, tab));
    SRED_TRY(ctx->misc.ensure(kMiscBytes));  // normalisation scratch, before capture
    SRED_HIP_CHECK(hipStreamSynchronize(ctx->stream));

    // ---- one graph: contractions, reductions, finalize
    hipStream_t cs = nullptr;
    SRED_HIP_CHECK(hipStreamCreate(&cs));
    hipStream_t saved = ctx->stream;
    ctx->stream = cs;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    int rc = SRED_OK;
    auto fail = [&](int code) {
        if (rc == SRED_OK) rc = code;
    };
    if (hipStreamBeginCapture(cs, hipStreamCaptureModeRelaxed) != hipSuccess) {
        set_error("hipStreamBeginCapture failed");
        fail(SRED_ERR_HIP);
    }
    const EpiConst ec = epi_const();
    int64_t pairs = 0;
    if (!rc) {
        int e = zero_dev(ctx, batch_counter(ctx) + 1, sizeof(unsigned long long));  // pairs skipped, all launches
        if (e) fail(e);
    }
    for (const Launch &L : launches) {
        if (rc) break;
        int e = launch_sym_metric(metric, L.KS, L.nitems, cs, ctx->batch_tab.as<SymProblem>() + L.tab_off, L.nprob,
                                  batch_counter(ctx), L.P, ec);
        if (e) { fail(e); break; }
        pairs += L.nitems * (int64_t)L.P * L.P;
        for (int i : L.items) {
            Item &x = it[i];
            e = global_reduce_dev(ctx, x.r, metric, d_part + x.part_off, x.npanels, x.Nv_pad, x.st->stds.as<float>(),
                                  x.st->S.as<float>());
            if (!e)
                e = global_finalize_dev(ctx, x.g, metric, x.st->S.as<float>(), x.st->W.as<float>(),
                                        x.st->mask.as<float>(), x.n_pixels, c, normalize, x.d_out);
            if (e) { fail(e); break; }
        }
    }
    for (int i = 0; i < n && !rc; i++)
        if (it[i].r.Nv == 0) {
            Item &x = it[i];
            int e = global_finalize_dev(ctx, x.g, metric, x.st->S.as<float>(), x.st->W.as<float>(),
                                        x.st->mask.as<float>(), x.n_pixels, c, normalize, x.d_out);
            if (e) fail(e);
        }
    hipError_t ce = hipStreamEndCapture(cs, &graph);
    if (ce != hipSuccess && !rc) {
        set_error(std::string("hipStreamEndCapture: ") + hipGetErrorString(ce));
        fail(SRED_ERR_HIP);
    }
    if (!rc && hipGraphInstantiate(&exec, graph, nullptr, nullptr, 0) != hipSuccess) {
        set_error("hipGraphInstantiate failed");
        fail(SRED_ERR_HIP);
    }
    if (!rc) {
        int e = time_begin(ctx);
        if (!e && hipGraphLaunch(exec, cs) != hipSuccess) {
            set_error("hipGraphLaunch failed");
            e = SRED_ERR_HIP;
        }
        if (!e) e = time_end(ctx, pairs);
        if (!e) ctx->last_skip = batch_counter(ctx) + 1;
        if (!e && hipStreamSynchronize(cs) != hipSuccess) {
            set_error("batch graph execution failed");
            e = SRED_ERR_HIP;
        }
        if (e) fail(e);
    }
    if (exec) (void)hipGraphExecDestroy(exec);
    if (graph) (void)hipGraphDestroy(graph);
    ctx->stream = saved;
    (void)hipStreamDestroy(cs);
    return rc;
}

}  // namespace sred
