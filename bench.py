"""Benchmark: global Pearson repetition map, 1024x1024 image, 11x11 blocks.

One step = the whole global-repetition pipeline on an image already in HBM
(local statistics, noise variance, relevance mask, compaction, block gather,
split-BF16 MFMA contraction with the fused epilogue, FP64 row reduction,
finalize and normalisation): sred_global_map_2d_dev.

Unit: one ordered (reference block, test block) pair over the interior,
N^2 with N = (1024-10)^2 = 1,028,196; value = N^2 / t / 1e9 (Gblock-pairs/s).
The split-BF16 contraction evaluates each unordered pair once (the per-pair
terms are symmetric) and skips pairs whose Gaussian weight is below 2^-40
(Pearson: t ranges more than sqrt(40) apart; SSIM: sqrt(47)), recognised per
panel pair and tile after sorting the blocks by sigma; the skipped count is
reported.  The roofline counts the algorithmic work of the reference's
evaluation, N^2 * 2K flops per map, over the contraction's time (HIP events on
its stream); the TFLOP/s the matrix cores issue and their busy fraction
(rocprofv3 counters, committed summary) are reported beside it.

The JSON line also carries: the plain-FP32 form timed the same way and the
speed-up; the split-BF16 form over all ordered pairs (the FP32 form's
algorithm) as all_pairs_value, for a like-for-like ratio; the SSIM and NRMSE
maps of the same image in both precisions; max/mean abs error of every
normalised map against the FP64 oracle on sampled pixels (plus the map
extremes) and between the two forms over the whole map; and the oracle's
float32 restatement of the reference timed on host cores on a bounded sample.

    python bench.py [--gpus 1] [--steps 5] [--warmup 1]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

W = H = 1024
B = 11
PEAK_BF16_TFLOPS = 2500.0   # MI355X dense BF16
PEAK_HBM_GBS = 8000.0
PEAK_FP32_VEC_TFLOPS = 157.3  # MI355X FP32 vector (packed FMA counted)
PEAK_FP64_VEC_TFLOPS = 78.6   # MI355X FP64 vector


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--fp32-steps", type=int, default=2)
    ap.add_argument("--oracle-rows", type=int, default=96)
    ap.add_argument("--cpu-seconds", type=float, default=12.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--profile", action="store_true",
                    help="short run for rocprofv3: bf16x3 steps only, then a 256 MiB device copy (PMC calibration)")
    ap.add_argument("--profile-block", type=int, default=B,
                    help="block side of the Pearson map a --profile run evaluates (odd, default 11)")
    ap.add_argument("--workload", choices=["global", "block", "adjoint", "batch", "hu", "phase", "vst", "quadtree",
                                           "timelapse", "tools", "vol", "block3d", "octree", "vst3d"],
                    default="global",
                    help="global: the headline metric (default); block: block repetition 2048x2048 with a 15x15 "
                         "ring template (four metrics and the Hu map); adjoint: global Pearson map + adjoint, "
                         "512x512, 9x9; batch: 64 images of sides U{128..1024} x block sizes {5,9,13,17} "
                         "in one batched call; hu: global Hu map 1024x1024, 11x11; phase: phase correlation map "
                         "256x256, 5x5; vst: variance-stabilised input image 1024x1024 (GAT fit, 5000 simplex "
                         "iterations, and transform); quadtree: the same image stabilised by the Quad/Octree "
                         "estimate; timelapse: block repetition over a 100-frame 512x512 stack, 15x15 ring template; "
                         "tools: the relevance-mask tools on a 1024x1024 image; vol: global Pearson map of a "
                         "128x128x64 volume, 5x5x5 blocks; block3d: block repetition of a 5x5x5 block over the "
                         "same volume (four metrics); octree: that volume's Poisson counts stabilised by the "
                         "Quad/Octree estimate; vst3d: the same counts through the VarianceStabilisingTransform3D_ "
                         "plugin (GATMinimizer3D, 5000 iterations)")
    return ap.parse_args()


def _timed(step, steps, warmup, kernel_ms):
    """Wall time per step (synchronised) and the mean dominant-kernel time."""
    import torch
    for _ in range(warmup):
        step()
    torch.cuda.synchronize()
    kern = []
    t0 = time.perf_counter()
    for _ in range(steps):
        step()
        kern.append(kernel_ms())
    torch.cuda.synchronize()
    return (time.perf_counter() - t0) / steps, float(np.mean(kern))


def block_run(args):
    """Block repetition, 2048x2048 rings-and-blobs image (3000 rings, 3000 blobs,
    Poisson noise at 200 photons), 15x15 ring template (K = 149): one map per
    metric through the whole block pipeline on device buffers, plus the Hu map.
    Roofline: the algorithmic HBM traffic of a block map, 4 B read + 4 B
    written per pixel, over the stencil kernel's time against 8 TB/s.  What
    bounds the kernel in practice is the vector ALU: the reference's
    sequential float sums per window (mean: K adds; variance: K x (sub, mul,
    add); template sum: K x (mul, add) on the same differences) are 6K FP32
    flops per pixel that cannot be reassociated; their rate is reported
    against the FP32 vector peak (157.3 TFLOP/s) and its half (78.6, the rate
    of one FMA per lane per cycle)."""
    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L
    from sred_amd import synthetic, utils
    from sred_amd.commands import block_map_device

    S, BB = 2048, 15
    K = sred_amd.block_size_2d(BB, BB)
    img = O.normalize_input(synthetic.rings_and_blobs(S, seed=1))
    ring = synthetic.ring_template(BB)
    ref_block = utils.get_reference_block_2d(ring)
    hu_tmpl = sred_amd.get_hu_template_2d(ring)
    d_img = torch.from_numpy(img).cuda()
    d_out = torch.zeros_like(d_img)
    ctx = sred_amd.Context(0)
    kernel_ms = lambda: ctx.last_contraction()[0]
    maps = {}
    for name in list(sred_amd.BLOCK_METRICS) + ["Hu"]:
        if name == "Hu":
            step = lambda: L.check(L.load().sred_block_hu_2d_dev(ctx.handle, d_img.data_ptr(), S, S,
                                                                L.fptr(L.f32(hu_tmpl).reshape(-1)), BB, BB,
                                                                d_out.data_ptr()))
        else:
            mid = L.load().sred_metric_from_name(name.encode(), 1)
            step = lambda mid=mid: block_map_device(ctx, mid, d_img.data_ptr(), S, S, ref_block, 0.0, True,
                                                    d_out.data_ptr())
        t, k = _timed(step, args.steps, args.warmup, kernel_ms)
        maps[name] = {"ms_per_map": t * 1e3, "kernel_ms": k, "Mpixel_per_s": S * S / t / 1e6,
                      "kernel_GB_per_s": 8.0 * S * S / (k * 1e-3) / 1e9,
                      "kernel_TFLOP_per_s": 6.0 * K * S * S / (k * 1e-3) / 1e12}
        out = d_out.cpu().numpy()
        if name == "Pearson's R":
            ref = O.block_map(O.PEARSON, img, ref_block.pixels, BB, BB, ref_block.mean, ref_block.std)
            maps[name]["max_abs_vs_refsim"] = float(np.abs(out - ref).max())
        print(f"# {name}: {t * 1e3:.3f} ms/map, kernel {k:.3f} ms", file=sys.stderr, flush=True)
    ctx.close()
    cpu = None
    if not args.no_cpu:
        cores = host_cores()
        t0 = time.perf_counter()
        O.block_map(O.PEARSON, img, ref_block.pixels, BB, BB, ref_block.mean, ref_block.std)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": S * S / t_cpu / 1e6, "unit": "Mpixel/s", "cores": cores, "kind": "port",
               "sample": f"oracle float32 restatement of kernelGetBlockPearson2D.cl + host steps, the full "
                         f"{S}x{S} Pearson map, {t_cpu:.2f} s"}
    p = maps["Pearson's R"]
    print(json.dumps({
        "metric": "Mpixel/s, block repetition Pearson map 2048² 15x15 ring template",
        "value": p["Mpixel_per_s"], "unit": "Mpixel/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": p["ms_per_map"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: 3000 rotated rings + 3000 Gaussian blobs, Poisson noise (200 photons), seed 1",
        "config": {"workload": "block repetition, 2048x2048, 15x15 ring template (K=149), relevance 0, normalised"},
        "maps": maps,
        "roofline": {"bound": "hbm", "achieved": p["kernel_GB_per_s"], "peak": PEAK_HBM_GBS, "unit": "GB/s",
                     "frac": p["kernel_GB_per_s"] / PEAK_HBM_GBS, "traffic": None, "bytes_per_pixel": 8,
                     "valu_tflops": p["kernel_TFLOP_per_s"], "flops_per_pixel": 6 * K,
                     "valu_frac_of_157_3": p["kernel_TFLOP_per_s"] / PEAK_FP32_VEC_TFLOPS,
                     "valu_frac_of_78_6": p["kernel_TFLOP_per_s"] / (PEAK_FP32_VEC_TFLOPS / 2),
                     "kernel": "k_block<Pearson>"},
        "cpu_baseline": cpu}))


def batch_run(args):
    """Batched global Pearson maps: 64 spot/ring mixture images (sides drawn
    from U{128..1024}, seed 3; image k uses seed 3+k), each at block sizes 5,
    9, 13 and 17 (256 problems), in one sred_global_batch_2d call: per-image
    front steps, then one graph launch with a persistent split-BF16 kernel per
    block-size class.  The first call (untimed) runs the panel autotuner."""
    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L
    from sred_amd import synthetic
    from sred_amd.commands import global_batch_device

    rng = np.random.default_rng(3)
    sides = rng.integers(128, 1025, size=(64, 2))
    imgs = [O.normalize_input(synthetic.spot_ring_mixture(int(w), int(h), seed=3 + k))
            for k, (w, h) in enumerate(sides)]
    blocks = (5, 9, 13, 17)
    d_imgs = [torch.from_numpy(im).cuda() for im in imgs]
    outs, probs, pairs, flops = [], [], 0.0, 0.0
    for im, d in zip(imgs, d_imgs):
        for b in blocks:
            o = torch.zeros_like(d)
            outs.append((im, b, o))
            probs.append((d.data_ptr(), im.shape[1], im.shape[0], b, b, o.data_ptr()))
            n = float((im.shape[0] - b + 1) * (im.shape[1] - b + 1))
            pairs += n * n
            flops += n * n * 2 * sred_amd.block_size_2d(b, b)
    ctx = sred_amd.Context(0)
    t0 = time.perf_counter()
    global_batch_device(ctx, L.PEARSON, probs)
    torch.cuda.synchronize()
    t_first = time.perf_counter() - t0
    print(f"# first call (with autotuning): {t_first:.2f} s", file=sys.stderr, flush=True)
    t, k = _timed(lambda: global_batch_device(ctx, L.PEARSON, probs), args.steps, 0,
                  lambda: ctx.last_contraction()[0])
    evaluated = ctx.last_contraction()[2]
    skipped = float(ctx.last_skipped())
    kss = sorted({(sred_amd.block_size_2d(b, b) + 31) // 32 for b in blocks})
    panels = {f"K_pad={32 * ks}": ctx.tuned_panel(L.PEARSON, ks) for ks in kss}
    tuning = {f"K_pad={32 * ks}": {str(P): ms for P, ms in ctx.tuning_ms(L.PEARSON, ks).items()} for ks in kss}
    ctx.close()
    # parity of every problem against the FP64 oracle: 16 seeded rows plus each
    # map's 4 smallest and 4 largest values (which fix its normalisation)
    t0 = time.perf_counter()
    worst, worst_at = 0.0, None
    for pi, (im, b, o) in enumerate(outs):
        out = o.cpu().numpy()
        r = b // 2
        m32, s32 = O.local_stats_f32(im, b, b)
        nv = O.noise_var(im)
        mask, thr, npx = O.relevance_mask(s32, r, r, nv, 0.0)
        idx = np.flatnonzero(mask.ravel() > 0)
        order = np.argsort(out.ravel()[idx], kind="stable")
        px = np.unique(np.concatenate([np.random.default_rng(pi).choice(idx, min(16, idx.size), replace=False),
                                       idx[order[:4]], idx[order[-4:]]]))
        S, Wt = O.global_accum_f64(O.PEARSON, im, s32, b, b, nv, thr, rows=px)
        sub = np.zeros_like(mask)
        sub.ravel()[px] = 1.0
        ref = O.global_finalize_f64(O.PEARSON, S, Wt, sub, npx, 0.0, True, b, b).ravel()[px]
        e = float(np.abs(out.ravel()[px] - ref).max())
        if e > worst:
            worst, worst_at = e, f"{im.shape[1]}x{im.shape[0]} b{b}"
    print(f"# every problem vs FP64 on sampled rows: max {worst:.2e} ({worst_at}), {time.perf_counter() - t0:.1f} s",
          file=sys.stderr, flush=True)
    # CPU baseline: the oracle's float32 restatement, whole rows of the largest
    # b = 17 problem against all of its blocks, on host cores
    cpu = None
    if not args.no_cpu:
        cores = host_cores()
        big = max((x for x in outs if x[1] == 17), key=lambda x: x[0].size)
        im, b = big[0], big[1]
        m32, s32 = O.local_stats_f32(im, b, b)
        nv = O.noise_var(im)
        r = b // 2
        inner = np.zeros(im.shape, bool)
        inner[r:im.shape[0] - r, r:im.shape[1] - r] = True
        rows_all = np.flatnonzero(inner)
        nb = float(rows_all.size)
        rows_done, t_cpu, step = 0, 0.0, 16
        while t_cpu < args.cpu_seconds and rows_done < rows_all.size:
            rows = rows_all[rows_done:rows_done + step]
            t1 = time.perf_counter()
            O.global_accum_f32(O.PEARSON, im, m32, s32, b, b, nv, 0.0, rows=rows, nthreads=cores)
            t_cpu += time.perf_counter() - t1
            rows_done += rows.size
            step = min(step * 2, 2048)
        cpu = {"value": rows_done * nb / t_cpu / 1e9, "unit": "Gblock-pairs/s", "cores": cores, "kind": "port",
               "sample": f"oracle float32 restatement of kernelGetGlobalPearson2D.cl on the largest 17x17 problem "
                         f"({im.shape[1]}x{im.shape[0]}): {rows_done} reference pixels x all {int(nb)} blocks, "
                         f"{t_cpu:.1f} s on {cores} threads"}
        print(f"# cpu: {cpu['value']:.4f} Gpairs/s", file=sys.stderr, flush=True)
    # algorithmic flops: every ordered pair of every problem, 2K each, over the graph time
    ach = flops / (k * 1e-3) / 1e12
    print(f"# batch: {t:.2f} s per call, graph {k / 1e3:.2f} s, {pairs:.3e} pairs, {skipped / pairs:.1%} skipped",
          file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "Gblock-pairs/s, batched global Pearson maps, 64 images x 4 block sizes",
        "value": pairs / t / 1e9, "unit": "Gblock-pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": 1,
        "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3",
        "data": "synthetic: spot/ring mixtures, sides U{128..1024} (seed 3), image k seed 3+k",
        "config": {"workload": "256 global Pearson maps (64 images x blocks 5,9,13,17), relevance 0, normalised, "
                               "one sred_global_batch_2d call", "pairs": pairs, "algorithmic_flops": flops},
        "graph_ms": k, "first_call_s": t_first, "tuned_panels": panels, "tuning_ms": tuning,
        "pairs_evaluated": evaluated, "pairs_skipped": skipped, "skip_fraction": skipped / pairs,
        "max_abs_vs_fp64_every_problem": worst, "worst_problem": worst_at,
        "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_BF16_TFLOPS, "traffic": None, "kernel": "graph (k_contract_mfma_sym x 4)"},
        "cpu_baseline": cpu}))


def pairmaps_run(args):
    """All-pairs vector maps (OpenCL kernels of the reference without a caller).

    hu:    global Hu map (kernelGetHuMap.cl), 1024x1024 two-texture image
           (ring lattice | blobs, seed 2), 11x11 blocks: every interior pixel
           against every interior pixel, 1.057e12 ordered pairs.  Algorithmic
           work 12 FP32 flops per pair (|ds|, sum, reciprocal, scale, exp2,
           |dhu|, fused accumulate), against the FP32 vector peak.
    phase: phase correlation map (textbook specification of
           kernelGetPhaseCorrelationMap.cl), 256x256 Gaussian-spot image
           (200 spots, seed 0), 5x5 blocks: 4.03e9 ordered pairs.  Algorithmic
           work per pair: the half-plane cross power (6 flops x 15
           frequencies), the column pass (8 flops x 5 x 5 x 2 minus the
           unit-twiddle terms) and the row pass (4 flops x 5 x 5 x 4 minus
           unit twiddles) = 790 FP64 flops, against the FP64 vector peak.
    One step = the whole map from the image and its local statistics in HBM
    (invariants or spectra, pair kernel, row reduction)."""
    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L
    from sred_amd import synthetic, utils

    hu = args.workload == "hu"
    if hu:
        S, BB, flops_pair = 1024, 11, 12.0
        img = O.normalize_input(synthetic.lattice_and_blobs(S, seed=2))
        fn = L.load().sred_global_hu_2d_dev
    else:
        S, BB, flops_pair = 256, 5, 790.0
        img = O.normalize_input(synthetic.spots(S, S, 200, seed=0))
        fn = L.load().sred_phase_correlation_2d_dev
    ctx = sred_amd.Context(0)
    st = sred_amd.get_local_statistics_2d(ctx, utils.InputImage2D(img, S, S, img.size), BB, BB)
    means, stds = st.local_means.reshape(S, S), st.local_stds.reshape(S, S)
    d_img, d_m, d_s = (torch.from_numpy(np.ascontiguousarray(a)).cuda() for a in (img, means, stds))
    d_out = torch.zeros_like(d_img)
    n = float((S - BB + 1) ** 2)
    pairs = n * n

    def step():
        L.check(fn(ctx.handle, d_img.data_ptr(), d_m.data_ptr(), d_s.data_ptr(), S, S, BB, BB, None, 0,
                   d_out.data_ptr()))

    t, k = _timed(step, args.steps, args.warmup, lambda: ctx.last_contraction()[0])
    ctx.close()
    out = d_out.cpu().numpy()
    rng = np.random.default_rng(7)
    r = BB // 2
    rows = (rng.integers(r, S - r, 8) * S + rng.integers(r, S - r, 8)).astype(np.int32)
    ref = (O.global_hu_map if hu else O.phase_map)(img, means, stds, BB, BB, rows=rows).reshape(-1)[rows]
    rel = float(np.max(np.abs(out.reshape(-1)[rows] - ref) / np.abs(ref)))
    cpu = None
    if not args.no_cpu:
        cores = host_cores()
        nr = 16 if hu else 48
        rows_c = (rng.integers(r, S - r, nr) * S + rng.integers(r, S - r, nr)).astype(np.int32)
        t0 = time.perf_counter()
        (O.global_hu_map if hu else O.phase_map)(img, means, stds, BB, BB, rows=rows_c)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": nr * n / t_cpu / 1e9, "unit": "Gpairs/s", "cores": cores, "kind": "port",
               "sample": f"oracle FP64 restatement, {nr} reference pixels x {int(n)} interior pixels, "
                         f"{t_cpu:.2f} s ({'hu invariants of the whole image included' if hu else 'spectra of the whole image included'})"}
    ach = pairs * flops_pair / (k * 1e-3) / 1e12
    peak = PEAK_FP32_VEC_TFLOPS if hu else PEAK_FP64_VEC_TFLOPS
    name = "global Hu map 1024² 11x11" if hu else "phase correlation map 256² 5x5"
    print(f"# {name}: {t * 1e3:.2f} ms/map, pair kernel {k:.2f} ms, max rel err (8 rows) {rel:.2e}",
          file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": f"Gpairs/s, {name}", "value": pairs / t / 1e9, "unit": "Gpairs/s", "n_gpus": 1,
        "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32" if hu else "f64",
        "data": ("synthetic: ring lattice (pitch 24) | Gaussian blobs, noise 0.05, seed 2" if hu
                 else "synthetic: 200 Gaussian spots (sigma 2), background 0.1, noise 0.05, seed 0"),
        "config": {"workload": name + ", every interior pixel against every interior pixel", "pairs": pairs},
        "kernel_ms": k, "max_rel_err_vs_oracle_8_rows": rel,
        "roofline": {"bound": "valu", "achieved": ach, "peak": peak, "unit": "TFLOP/s", "frac": ach / peak,
                     "traffic": None,
                     "kernel": ("k_hu_pairs_sym" if hu else "k_phase_pairs_sym<5, 5, true>")
                               + " (each unordered pair once)",
                     "issued_tflops": ach / 2.0,
                     "flops_per_pair": flops_pair},
        "cpu_baseline": cpu}))


def vol_run(args):
    """GlobalRepetition3D_ with Pearson's R (as intended where the reference
    does not run) on a 128x128x64 volume of Gaussian spheres with noise
    (seed 5), 5x5x5 blocks (ellipsoid K = 33, padded to 64), relevance 0,
    normalised, split-BF16 contraction; the volume in HBM.  Same metric and
    roofline as the 2D headline: Gblock-pairs/s over all ordered pairs of
    inner voxels, the MFMA roofline of the symmetric contraction at the
    algorithmic 2K flops per evaluated pair."""
    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L

    D, H, W, B = 64, 128, 128, 5
    blobs = _spheres
    vol = blobs(D, H, W, 5, 2000)
    d_vol = torch.from_numpy(vol).cuda()
    d_out = torch.empty_like(d_vol)
    ctx = sred_amd.Context(0)
    lib = L.load()

    def step():
        L.check(lib.sred_global_repetition_3d_dev(ctx.handle, L.PEARSON, L.BF16X3, d_vol.data_ptr(), W, H, D, B, B, B,
                                                  0.0, 1, d_out.data_ptr(), None, None, None, None, None))

    t, kms = _timed(step, args.steps, args.warmup, lambda: ctx.last_contraction()[0])
    _, _, evaluated = ctx.last_contraction()
    skipped = ctx.last_skipped()
    ctx.close()
    inner = (W - 2 * (B // 2)) * (H - 2 * (B // 2)) * (D - 2 * (B // 2))
    pairs = float(inner) * inner
    K = O.footprint3d(B, B, B)[0]
    ach = (pairs - skipped) * 2 * K / (kms * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        small = blobs(24, 40, 40, 6, 40)
        t0 = time.perf_counter()
        O.global_pearson3d(small, B, B, B)
        t_cpu = time.perf_counter() - t0
        n_small = (40 - 4) * (40 - 4) * (24 - 4)
        cpu = {"value": n_small * n_small / t_cpu / 1e9, "unit": "Gblock-pairs/s", "cores": host_cores(),
               "kind": "port", "sample": f"oracle float restatement (OpenMP) of the whole map of a 40x40x24 volume, "
                                         f"5x5x5 blocks, {t_cpu:.1f} s"}
    print(f"# vol: {t * 1e3:.1f} ms per map, contraction {kms:.1f} ms, {skipped / pairs:.1%} of pairs skipped",
          file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "Gblock-pairs/s, global Pearson map of a 128x128x64 volume, 5x5x5 blocks",
        "value": pairs / t / 1e9, "unit": "Gblock-pairs/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "bf16x3",
        "data": "synthetic: 2000 Gaussian spheres (sigma 2) plus noise in a 128x128x64 volume, seed 5",
        "config": {"workload": f"GlobalRepetition3D_ Pearson, 128x128x64, 5x5x5 (K={K}), relevance 0, normalised"},
        "contraction_ms": kms, "pairs_skipped": float(skipped), "skip_fraction": skipped / pairs,
        "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_BF16_TFLOPS, "traffic": None, "flops_per_pair": 2 * K},
        "cpu_baseline": cpu}))


def _spheres(d, h, w, seed, n):
    """Gaussian spheres (sigma 2, amplitudes U(0.5, 1)) on 0.1 plus N(0, 0.03) noise, min-max normalised."""
    rng = np.random.default_rng(seed)
    v = np.full((d, h, w), 0.1, np.float32)
    r = 7  # each sphere within +-3.5 sigma
    for _ in range(n):
        cz, cy, cx = rng.uniform(0, d), rng.uniform(0, h), rng.uniform(0, w)
        a = np.float32(rng.uniform(0.5, 1.0))
        z0, z1 = max(0, int(cz) - r), min(d, int(cz) + r + 1)
        y0, y1 = max(0, int(cy) - r), min(h, int(cy) + r + 1)
        x0, x1 = max(0, int(cx) - r), min(w, int(cx) + r + 1)
        zz, yy, xx = np.ogrid[z0:z1, y0:y1, x0:x1]
        v[z0:z1, y0:y1, x0:x1] += a * np.exp(
            -((zz - cz) ** 2 + (yy - cy) ** 2 + (xx - cx) ** 2) / (2 * 2.0 ** 2)).astype(np.float32)
    v += np.float32(0.03) * rng.standard_normal(v.shape).astype(np.float32)
    return ((v - v.min()) / (v.max() - v.min())).astype(np.float32)


def block3d_run(args):
    """BlockRepetition3D_ (as intended) on the 128x128x64 sphere volume of the
    vol workload (seed 5, normalised, in HBM) with a 5x5x5 reference block
    (K = 33) cut around voxel (30, 60, 70) with N(0, 0.02) noise: one map per
    metric through the whole pipeline (statistics, kernel, normalisation).
    Headline: the Pearson map in Mvoxel/s.  Roofline of the k_block3d
    kernel: the vector ALU at the reference's per-voxel arithmetic for
    Pearson, 11K operations (per element: two min/max pairs, two remaps of
    a subtraction and a division, the local-mean subtraction, a multiply-add;
    each division counted as one operation, though it is a double division
    of about ten FP64 instructions) against the FP32 vector peak."""
    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L

    D, H, W, B = 64, 128, 128, 5
    vol = _spheres(D, H, W, 5, 2000)
    blk = vol[28:33, 58:63, 68:73] + np.random.default_rng(7).normal(0.0, 0.02, (B, B, B)).astype(np.float32)
    ref = sred_amd.get_reference_block_3d(blk)
    K = ref.size
    d_vol = torch.from_numpy(vol).cuda()
    d_out = torch.empty_like(d_vol)
    ctx = sred_amd.Context(0)
    lib = L.load()
    tp = L.f32(ref.pixels)
    inner = (W - 2 * (B // 2)) * (H - 2 * (B // 2)) * (D - 2 * (B // 2))
    maps = {}
    for name in sred_amd.BLOCK_METRICS:
        mid = lib.sred_metric_from_name(name.encode(), 1)

        def step(mid=mid):
            L.check(lib.sred_block_repetition_3d_dev(ctx.handle, mid, d_vol.data_ptr(), W, H, D, L.fptr(tp), B, B, B,
                                                     ref.mean, ref.std, 0.0, 1, d_out.data_ptr(), None, None, None))

        t, k = _timed(step, args.steps, args.warmup, lambda: ctx.last_contraction()[0])
        maps[name] = {"ms_per_map": t * 1e3, "kernel_ms": k, "Mvoxel_per_s": D * H * W / t / 1e6}
        if name == "Pearson's R":
            want = O.block_map3d(O.PEARSON, vol, ref.pixels, B, B, B, ref.mean, ref.std)
            maps[name]["max_abs_vs_refsim"] = float(np.abs(d_out.cpu().numpy() - want).max())
        print(f"# block3d {name}: {t * 1e3:.3f} ms/map, kernel {k:.3f} ms", file=sys.stderr, flush=True)
    ctx.close()
    p = maps["Pearson's R"]
    ach = 11.0 * K * inner / (p["kernel_ms"] * 1e-3) / 1e12
    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        O.block_map3d(O.PEARSON, vol, ref.pixels, B, B, B, ref.mean, ref.std)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": D * H * W / t_cpu / 1e6, "unit": "Mvoxel/s", "cores": host_cores(), "kind": "port",
               "sample": f"oracle float restatement (OpenMP) of the whole Pearson map, {t_cpu:.2f} s"}
    print(json.dumps({
        "metric": "Mvoxel/s, block repetition Pearson map of a 128x128x64 volume, 5x5x5 block",
        "value": p["Mvoxel_per_s"], "unit": "Mvoxel/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": p["ms_per_map"], "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32", "data": "synthetic: 2000 Gaussian spheres (sigma 2) plus noise in a 128x128x64 volume, "
                                "seed 5; block cut at (30, 60, 70) plus N(0, 0.02), seed 7",
        "config": {"workload": f"BlockRepetition3D_, 128x128x64, 5x5x5 block (K={K}), relevance 0, normalised"},
        "maps": maps,
        "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_VEC_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP32_VEC_TFLOPS, "traffic": None, "ops_per_voxel": 11 * K,
                     "kernel": "k_block3d<Pearson>"},
        "cpu_baseline": cpu}))


def tools_run(args):
    """The relevance-mask tools on a 1024x1024 Poisson rings-and-blobs image
    (seed 1), 3x3 blocks: the shared input step (GAT fit from gain 0, 5000
    iterations at most, GAT, double normalisation), then RelevanceMask2D_'s
    mask (constant 1.0) and OptimiseRelevanceMask_'s scan (step 0.001) on the
    prepared image, device buffers throughout.  One step = the input step
    plus both tools; their times are reported separately.  The input step is
    the simplex fit (see the vst workload); the tools are a stencil, a tile
    reduction and, for the scan, one extremes reduction and a host loop."""
    import ctypes

    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L
    from sred_amd import synthetic

    S = 1024
    img = synthetic.rings_and_blobs(S, seed=1)
    d_src = torch.from_numpy(img).cuda()
    d_img = torch.empty_like(d_src)
    d_mask = torch.empty_like(d_src)
    ctx = sred_amd.Context(0)
    p = (ctypes.c_double * 3)()
    nv, thr, ones, zeros = ctypes.c_float(), ctypes.c_float(), ctypes.c_float(), ctypes.c_float()
    lib = L.load()
    parts = {"prep": [], "mask": [], "scan": []}

    def step():
        t0 = time.perf_counter()
        d_img.copy_(d_src)
        L.check(lib.sred_tool_prep_2d_dev(ctx.handle, d_img.data_ptr(), S, S, 5000, p))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        L.check(lib.sred_relevance_mask_tool_2d_dev(ctx.handle, d_img.data_ptr(), S, S, 3, 3, 1.0, d_mask.data_ptr(),
                                                    ctypes.byref(nv), ctypes.byref(thr)))
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        L.check(lib.sred_optimise_relevance_mask_2d_dev(ctx.handle, d_img.data_ptr(), S, S, 3, 3, 0.001, 100000000,
                                                        ctypes.byref(ones), ctypes.byref(zeros)))
        t3 = time.perf_counter()
        parts["prep"].append(t1 - t0)
        parts["mask"].append(t2 - t1)
        parts["scan"].append(t3 - t2)

    t, _ = _timed(step, args.steps, args.warmup, lambda: 0.0)
    ctx.close()
    k = args.steps
    ms = {name: 1e3 * float(np.mean(v[-k:])) for name, v in parts.items()}
    cpu = None
    if not args.no_cpu:
        prep = d_img.cpu().numpy()
        t0 = time.perf_counter()
        O.relevance_mask_tool(prep, 3, 3, 1.0)
        O.optimise_relevance_mask(prep, 3, 3, 0.001)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": t_cpu * 1e3, "unit": "ms for both tools", "cores": 1, "kind": "port",
               "sample": f"oracle restatements of RelevanceMask2D_ and OptimiseRelevanceMask_ (the scan rebuilding "
                         f"every mask, as the reference does) on the prepared {S}x{S} image"}
    print(f"# tools: {t * 1e3:.2f} ms per step (input step {ms['prep']:.2f}, mask {ms['mask']:.3f}, scan "
          f"{ms['scan']:.3f} ms), constants ({ones.value}, {zeros.value})", file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "ms per relevance-mask tool run (input step + RelevanceMask2D_ + OptimiseRelevanceMask_), 1024²",
        "value": t * 1e3, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: Poisson counts of 3000 rings + 3000 blobs at 1024x1024 (200 photons peak), seed 1",
        "config": {"workload": "RelevanceMask2D_ (3x3, constant 1.0) and OptimiseRelevanceMask_ (3x3, step 0.001) "
                               "after their shared input step, 1024x1024"},
        "input_step_ms": ms["prep"], "mask_ms": ms["mask"], "scan_ms": ms["scan"],
        "gat_fit": list(p), "noise_var": nv.value, "constant_range": [ones.value, zeros.value],
        "roofline": None, "cpu_baseline": cpu}))


def timelapse_run(args):
    """The time-lapse branch of BlockRepetition2D_ (BlockRepetition2D_.java:95-137),
    no stabilisation, Pearson, relevance constant 0, normalised: a stack of
    100 frames of 512x512 Poisson rings and blobs (seed 1 + t), 15x15 ring
    template (K = 149), stack and maps resident in HBM.  One step = the whole
    stack: stack and frame normalisations, one block pipeline per frame, the
    final stack normalisation.  Roofline of the stencil: the FP32 vector ALU
    at 6K flops per pixel, as the single-image block workload."""
    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L
    from sred_amd import synthetic

    D, S, T = 100, 512, 15
    frames = [synthetic.rings_and_blobs(S, 180, 180, seed=1 + t) for t in range(D)]
    st = np.stack(frames).astype(np.float32)
    tmpl = synthetic.ring_template(T)
    rb = sred_amd.get_reference_block_2d(tmpl)
    d_st = torch.from_numpy(st).cuda()
    d_out = torch.empty_like(d_st)
    ctx = sred_amd.Context(0)
    tp = L.f32(rb.pixels)

    def step():
        L.check(L.load().sred_block_timelapse_2d_dev(ctx.handle, O.PEARSON, d_st.data_ptr(), S, S, D, L.fptr(tp), T,
                                                     T, rb.mean, rb.std, 0.0, 1, d_out.data_ptr()))

    t, kms = _timed(step, args.steps, args.warmup, lambda: ctx.last_contraction()[0])
    ctx.close()
    px = D * S * S
    flops = 6.0 * rb.size * px
    ach = flops / t / 1e12
    cpu = None
    if not args.no_cpu:
        n_fr = 30
        t0 = time.perf_counter()
        O.block_timelapse(O.PEARSON, st[:n_fr], rb.pixels, T, T, rb.mean, rb.std, 0.0, True)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": n_fr / t_cpu, "unit": "frames/s", "cores": host_cores(), "kind": "port",
               "sample": f"oracle restatement (refsim block maps, OpenMP), first {n_fr} frames of the stack, "
                         f"{t_cpu:.2f} s"}
    print(f"# timelapse: {t * 1e3:.2f} ms per {D}-frame stack ({t * 1e3 / D:.3f} ms per frame), stencil "
          f"{kms:.2f} ms", file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "frames per second, time-lapse block repetition (512², 15×15 ring template, Pearson)",
        "value": D / t, "unit": "frames/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": f"synthetic: {D} frames of Poisson rings and blobs at {S}x{S}, seeds 1..{D}",
        "config": {"workload": f"BlockRepetition2D_ time-lapse, {D} frames {S}x{S}, 15x15 ring template (K="
                               f"{rb.size}), relevance 0, normalised"},
        "stencil_ms": kms,
        "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_FP32_VEC_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP32_VEC_TFLOPS, "traffic": None,
                     "flops_per_pixel": 6 * rb.size, "note": "whole step, launch overheads included"},
        "cpu_baseline": cpu}))


def quadtree_run(args):
    """Variance-stabilised input image by the Quad/Octree estimate
    (Utils.getInputImage2D, Utils.java:686-703: QuadTree_(w, h, 4, 0.01f),
    50 robust-mean iterations, LTS fraction 0.75f, applyGATtree), 1024x1024
    Poisson counts of rings and blobs (200 photons at peak, seed 1), on the
    device image.  The dominant kernel, k_qt_criteria, evaluates every node
    that could divide; its time is the root's in-order float chain (the
    reference's sequential sums over all w*h pixels: 3 streams, then 1), so
    it is reported as a critical path (ns per chained element), not against
    a throughput roofline."""
    import ctypes

    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L
    from sred_amd import synthetic

    S = 1024
    img = synthetic.rings_and_blobs(S, seed=1)
    d_img = torch.from_numpy(img).cuda()
    d_work = torch.empty_like(d_img)
    ctx = sred_amd.Context(0)
    ge = (ctypes.c_double * 2)()
    nl = ctypes.c_int()

    def step():
        L.check(L.load().sred_quadtree_gat_2d_dev(ctx.handle, d_img.data_ptr(), S, S, 4, 0.01, 50, 0.75, ge,
                                                  ctypes.byref(nl), 0, None, None, None))
        d_work.copy_(d_img)
        L.check(L.load().sred_gat_tree_2d_dev(ctx.handle, d_work.data_ptr(), S * S, ge[0], ge[1]))
        mn, mx = d_work.min(), d_work.max()
        d_work.sub_(mn).div_(mx - mn + 1e-7)

    t, kms = _timed(step, args.steps, args.warmup, lambda: ctx.last_contraction()[0])
    ctx.close()
    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        ref = O.quadtree_gat(img)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": t_cpu * 1e3, "unit": "ms per estimate", "cores": 1, "kind": "port",
               "sample": f"oracle restatement of QuadTree_ (same tree, float orders), the full {S}x{S} estimate "
                         f"({len(ref.means)} leaves)"}
        assert (ge[0], ge[1]) == (ref.g0, ref.edc), "device estimate differs from the oracle"
    chain = 2 * S * S  # the root's chained elements: pass 1 (3 streams interleaved) + pass 2
    print(f"# quadtree: {t * 1e3:.2f} ms per stabilised image, {nl.value} leaves, criteria kernel {kms:.2f} ms, "
          f"g0 {ge[0]:.6g} eDC {ge[1]:.6g}", file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "ms per variance-stabilised 1024² input image (Quad/Octree GAT estimate and transform)",
        "value": t * 1e3, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: Poisson counts of 3000 rings + 3000 blobs at 1024x1024 (200 photons peak), seed 1",
        "config": {"workload": "Utils.getInputImage2D, Quad/Octree GAT (min size 4, alpha 0.01, 50 iterations, "
                               "LTS 0.75), 1024x1024"},
        "leaves": nl.value, "g0": ge[0], "eDC": ge[1], "criteria_kernel_ms": kms,
        "roofline": None,
        "critical_path": {"chained_elements": chain, "ns_per_element": kms * 1e6 / chain,
                          "note": "the root node's sequential float sums (reference order) bound the criteria kernel"},
        "cpu_baseline": cpu}))


def octree_run(args):
    """Variance-stabilised input volume by the Quad/Octree estimate
    (Utils.getInputImage3D, Utils.java:1234-1260: OcTree_(w, h, d, 4, 0.01f),
    50 robust-mean iterations, LTS fraction 0.75f, applyGATtree over every
    voxel, normalisation): gain-2 Poisson counts of the vol workload's
    128x128x64 sphere volume (200 photons at the brightest, offset 10, read
    noise 3; seed 5), on the device volume.  As for the quadtree, the
    criteria kernel is bound by the root's in-order chains (double sums of v
    and v*v and the float residual sum over all voxels, then one float
    chain), reported as a critical path."""
    import ctypes

    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L

    D, H, W = 64, 128, 128
    rng = np.random.default_rng(5)
    vol = (2.0 * rng.poisson(200.0 * _spheres(D, H, W, 5, 2000)) + rng.normal(10.0, 3.0, (D, H, W))).astype(np.float32)
    d_vol = torch.from_numpy(vol).cuda()
    d_work = torch.empty_like(d_vol)
    ctx = sred_amd.Context(0)
    ge = (ctypes.c_double * 2)()
    nl = ctypes.c_int()
    lib = L.load()

    def step():
        L.check(lib.sred_octree_gat_3d_dev(ctx.handle, d_vol.data_ptr(), W, H, D, 4, 0.01, 50, 0.75, ge,
                                           ctypes.byref(nl), 0, None, None, None))
        d_work.copy_(d_vol)
        L.check(lib.sred_gat_tree_3d_dev(ctx.handle, d_work.data_ptr(), D * H * W, ge[0], ge[1]))
        mn, mx = d_work.min(), d_work.max()
        d_work.sub_(mn).div_(mx - mn + 1e-7)

    t, kms = _timed(step, args.steps, args.warmup, lambda: ctx.last_contraction()[0])
    ctx.close()
    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        ref = O.octree_gat(vol)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": t_cpu * 1e3, "unit": "ms per estimate", "cores": 1, "kind": "port",
               "sample": f"oracle restatement of OcTree_ (same tree, float and double orders), the full "
                         f"{W}x{H}x{D} estimate ({len(ref.means)} leaves)"}
        assert (ge[0], ge[1]) == (ref.g0, ref.edc), "device estimate differs from the oracle"
    chain = D * H * W  # the root's double chain (sums of v and v*v together); its two float chains run on another wave
    print(f"# octree: {t * 1e3:.2f} ms per stabilised volume, {nl.value} leaves, criteria kernel {kms:.2f} ms, "
          f"g0 {ge[0]:.6g} eDC {ge[1]:.6g}", file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "ms per variance-stabilised 128x128x64 input volume (Quad/Octree GAT estimate and transform)",
        "value": t * 1e3, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
        "data": "synthetic: gain-2 Poisson counts of 2000 Gaussian spheres (200 photons peak) + N(10, 3), "
                "128x128x64, seed 5",
        "config": {"workload": "Utils.getInputImage3D, Quad/Octree GAT (min size 4, alpha 0.01, 50 iterations, "
                               "LTS 0.75), 128x128x64"},
        "leaves": nl.value, "g0": ge[0], "eDC": ge[1], "criteria_kernel_ms": kms,
        "roofline": None,
        "critical_path": {"chained_elements": chain, "ns_per_element": kms * 1e6 / chain,
                          "note": "the root node's sequential double sums (reference order) bound the criteria "
                                  "kernel; its two float passes (2x as many adds) run beside them"},
        "cpu_baseline": cpu}))


def vst_run(args):
    """Variance-stabilised input image (Utils.getInputImage2D with the Simplex
    GAT, maxIter 5000 as GlobalRepetition2D_.java:114): 1024x1024 Poisson
    counts (rings and blobs, 200 photons at peak, seed 1).  One step = the
    simplex fit on the device image (each iteration one launch that evaluates
    the reflection, expansion and both contractions and writes the four
    values to pinned host memory) plus the transform and the min-max
    normalisation.  Roofline of the objective's row kernel: the FP64 vector
    ALU, 10 algorithmic flops per pixel and parameter set (g*v + ref, the
    square root, the scale by 2/g, the float square, two double adds, the
    comparison with -ref/g and the two conversions counted as one each)."""
    import ctypes

    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L
    from sred_amd import synthetic

    S, ITERS = 1024, 5000
    img = synthetic.rings_and_blobs(S, seed=1)
    d_img = torch.from_numpy(img).cuda()
    d_work = torch.empty_like(d_img)
    ctx = sred_amd.Context(0)
    params = (ctypes.c_double * 3)()
    err = ctypes.c_double()

    def step():
        L.check(L.load().sred_gat_minimize_2d_dev(ctx.handle, d_img.data_ptr(), S, S, 1.0, 10.0, 100.0, ITERS,
                                                  params, ctypes.byref(err)))
        d_work.copy_(d_img)
        g, s, o = (float(np.float32(v)) for v in params)
        L.check(L.load().sred_gat_2d_dev(ctx.handle, d_work.data_ptr(), S * S, g, s, o))
        mn, mx = d_work.min(), d_work.max()
        d_work.sub_(mn).div_(mx - mn + 1e-7)

    t, _ = _timed(step, args.steps, args.warmup, lambda: 0.0)
    # the row kernel's time: one more fit with HIP events around every launch
    # (kept out of the timed steps, since the events add about 4 us a launch)
    ctx.set_option(L.OPT_FIT_TIMING, 1)
    step()
    torch.cuda.synchronize()
    kms, launches, evals = ctx.last_contraction()
    ctx.set_option(L.OPT_FIT_TIMING, 0)
    fit = tuple(params)
    ctx.close()
    cpu = None
    if not args.no_cpu:
        cores = host_cores()
        n_it = 20
        t0 = time.perf_counter()
        O.gat_minimize(img, 1.0, 10.0, 100.0, n_it)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": n_it / t_cpu, "unit": "simplex iterations/s", "cores": 1, "kind": "port",
               "sample": f"oracle restatement (same simplex, same objective), {n_it} iterations of the {S}x{S} "
                         f"fit, {t_cpu:.2f} s"}
    flops = 10.0 * evals  # per pixel and parameter set, as above
    ach = flops / (kms * 1e-3) / 1e12
    print(f"# vst: {t * 1e3:.1f} ms per stabilised image, {launches} objective launches, fit {fit}",
          file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "ms per variance-stabilised 1024² input image (GAT fit, 5000 simplex iterations)",
        "value": t * 1e3, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: Poisson counts of 3000 rings + 3000 blobs at 1024x1024 (200 photons peak), seed 1",
        "config": {"workload": "Utils.getInputImage2D, Simplex GAT, maxIter 5000, 1024x1024"},
        "objective_launches": launches, "ms_per_launch": t * 1e3 / max(launches, 1), "fit": fit,
        "objective": err.value, "objective_pixel_evals_per_s": evals / t,
        "rows_kernel_ms_per_launch": kms / max(launches, 1),
        "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_FP64_VEC_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP64_VEC_TFLOPS, "traffic": None,
                     "flops_per_pixel_set": 10, "pixel_set_evals_per_launch": evals / max(launches, 1)},
        "cpu_baseline": cpu}))


def vst3d_run(args):
    """The VarianceStabilisingTransform3D_ plugin (VarianceStabilisingTransform3D_
    .java:28-110) on the octree workload's volume (gain-2 Poisson counts of
    the sphere volume, 128x128x64, seed 5): GATMinimizer3D from (gain, sigma,
    offset) = (1, 100, 10) with 5000 iterations, then the GAT with the fitted
    doubles, on the device volume.  Roofline of the objective's row kernel
    as for the 2D fit: the FP64 vector ALU at 10 algorithmic flops per voxel
    and parameter set (the square now in double)."""
    import ctypes

    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L

    D, H, W, ITERS = 64, 128, 128, 5000
    rng = np.random.default_rng(5)
    vol = (2.0 * rng.poisson(200.0 * _spheres(D, H, W, 5, 2000)) + rng.normal(10.0, 3.0, (D, H, W))).astype(np.float32)
    d_vol = torch.from_numpy(vol).cuda()
    d_work = torch.empty_like(d_vol)
    ctx = sred_amd.Context(0)
    params = (ctypes.c_double * 3)()
    err = ctypes.c_double()
    lib = L.load()

    def step():
        L.check(lib.sred_gat_minimize_3d_dev(ctx.handle, d_vol.data_ptr(), W, H, D, 1.0, 100.0, 10.0, ITERS, params,
                                             ctypes.byref(err)))
        d_work.copy_(d_vol)
        L.check(lib.sred_gat_2d_dev(ctx.handle, d_work.data_ptr(), D * H * W, params[0], params[1], params[2]))

    t, _ = _timed(step, args.steps, args.warmup, lambda: 0.0)
    ctx.set_option(L.OPT_FIT_TIMING, 1)
    step()
    torch.cuda.synchronize()
    kms, launches, evals = ctx.last_contraction()
    ctx.set_option(L.OPT_FIT_TIMING, 0)
    fit = tuple(params)
    ctx.close()
    cpu = None
    if not args.no_cpu:
        n_it = 20
        t0 = time.perf_counter()
        p, _ = O.gat_minimize3d(vol, 1.0, 100.0, 10.0, n_it)
        t_cpu = time.perf_counter() - t0
        cpu = {"value": n_it / t_cpu, "unit": "simplex iterations/s", "cores": 1, "kind": "port",
               "sample": f"oracle restatement (same simplex, same objective), {n_it} iterations of the "
                         f"{W}x{H}x{D} fit, {t_cpu:.2f} s"}
    ach = 10.0 * evals / (kms * 1e-3) / 1e12
    print(f"# vst3d: {t * 1e3:.1f} ms per stabilised volume, {launches} objective launches, fit {fit}",
          file=sys.stderr, flush=True)
    print(json.dumps({
        "metric": "ms per variance-stabilised 128x128x64 volume (GATMinimizer3D fit, 5000 simplex iterations)",
        "value": t * 1e3, "unit": "ms", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": t * 1e3, "higher_is_better": False, "scaling": "weak", "vs_baseline": None, "dtype": "f64",
        "data": "synthetic: gain-2 Poisson counts of 2000 Gaussian spheres (200 photons peak) + N(10, 3), "
                "128x128x64, seed 5",
        "config": {"workload": "VarianceStabilisingTransform3D_, GATMinimizer3D from (1, 100, 10), maxIter 5000, "
                               "128x128x64"},
        "objective_launches": launches, "ms_per_launch": t * 1e3 / max(launches, 1), "fit": fit,
        "objective": err.value, "rows_kernel_ms_per_launch": kms / max(launches, 1),
        "roofline": {"bound": "valu", "achieved": ach, "peak": PEAK_FP64_VEC_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_FP64_VEC_TFLOPS, "traffic": None,
                     "flops_per_voxel_set": 10, "voxel_set_evals_per_launch": evals / max(launches, 1)},
        "cpu_baseline": cpu}))


def adjoint_run(args):
    """Global Pearson map + hand-written adjoint, 512x512 spots image (800
    spots), 9x9 blocks (K = 49), loss weights U[0,1] (seed 40): one step =
    forward (split-BF16 symmetric contraction) and adjoint on device buffers.
    The adjoint kernel recomputes every dot product (2K flops per ordered
    pair) and accumulates the row gradient (2K more), both as split-BF16
    MFMA products (three bf16 products each; the dot products at K padded to
    64, the gradient product over k < 48 = K - 1, its last component following
    from the centring)."""
    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L
    from sred_amd import synthetic

    S, BB = 512, 9
    img = O.normalize_input(synthetic.spots(S, S, 800, seed=4))
    u = np.random.default_rng(40).random((S, S)).astype(np.float32)
    d_img = torch.from_numpy(img).cuda()
    d_u = torch.from_numpy(u).cuda()
    d_grad = torch.zeros_like(d_img)
    d_map = torch.zeros_like(d_img)
    ctx = sred_amd.Context(0)
    Kb = sred_amd.block_size_2d(BB, BB)
    N = (S - 2 * (BB // 2)) ** 2

    def step():
        L.check(L.load().sred_global_pearson_adjoint_2d_dev(ctx.handle, d_img.data_ptr(), S, S, BB, BB,
                                                           d_u.data_ptr(), d_grad.data_ptr(), d_map.data_ptr()))

    t, k = _timed(step, args.steps, args.warmup, lambda: ctx.last_contraction()[0])
    pairs_adj = ctx.last_contraction()[2]
    ctx.close()
    flops = float(pairs_adj) * 4 * Kb
    ach = flops / (k * 1e-3) / 1e12
    kpad = (Kb + 31) // 32 * 32
    kgrad = (Kb - 1 + 15) // 16 * 16
    issued = float(pairs_adj) * 2 * 3 * (kpad + kgrad) / (k * 1e-3) / 1e12
    g = d_grad.cpu().numpy()
    cpu = None
    if not args.no_cpu:
        # CPU baseline: the FP64 torch autograd restatement of the map and its
        # gradient (the reference has no adjoint), on a 120x120 crop of the
        # same image (N = 12,544 blocks: its N x N pair matrices take ~1.3 GB
        # each), forward + backward per call
        from oracle import torch_ref as T
        crop, uc = img[:120, :120], u[:120, :120]
        Nc = (120 - 2 * (BB // 2)) ** 2
        T.pearson_map_grad(crop, BB, uc.astype(np.float64))
        reps, t0 = 2, time.perf_counter()
        for _ in range(reps):
            T.pearson_map_grad(crop, BB, uc.astype(np.float64))
        tc = (time.perf_counter() - t0) / reps
        cpu = {"value": float(Nc) ** 2 / tc / 1e9, "unit": "Gblock-pairs/s", "cores": torch.get_num_threads(),
               "kind": "port", "sample": f"FP64 torch autograd restatement of the map and its gradient "
               f"(oracle/torch_ref.py) on a 120x120 crop of the same image and weights, 9x9 blocks, "
               f"{tc:.2f} s per forward + backward"}
        print(f"# cpu: {cpu['value']:.4f} Gpairs/s ({tc:.2f} s per call)", file=sys.stderr)
    print(json.dumps({
        "metric": "Gblock-pairs/s, global Pearson map + adjoint 512² b9", "value": float(N) ** 2 / t / 1e9,
        "unit": "Gblock-pairs/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": t * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "bf16x3",
        "data": "synthetic: 800 Gaussian spots, noise 0.05, seed 4; loss weights U[0,1] seed 40",
        "config": {"workload": "forward + adjoint of the normalised global Pearson map, 512x512, 9x9 (K=49)"},
        "adjoint_kernel_ms": k, "grad_abs_max": float(np.abs(g).max()),
        "roofline": {"bound": "mfma", "achieved": ach, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": ach / PEAK_BF16_TFLOPS, "traffic": None, "issued_tflops": issued,
                     "kernel": f"k_adjoint_mfma<KS={kpad // 32}, KG={kgrad // 16}>"},
        "cpu_baseline": cpu}))


def profile_run(args):
    """bf16x3 steps only plus one plain copy of known size, so rocprofv3 PMC
    counters (FETCH_SIZE, WRITE_SIZE) of the contraction can be scaled by
    what the same counters read for a copy of 256 MiB."""
    import torch

    import sred_amd
    from sred_amd import _lib as L
    from sred_amd import synthetic
    from sred_amd.commands import global_map_device

    bb = args.profile_block
    if bb < 3 or bb % 2 == 0:
        raise SystemExit("--profile-block must be odd and at least 3")
    img = sred_amd.get_input_image_2d(synthetic.lattice_and_blobs(W, seed=2)).image_array
    d_img = torch.from_numpy(img).cuda()
    d_out = torch.zeros_like(d_img)
    with sred_amd.Context(0) as ctx:
        for i in range(args.warmup + args.steps):
            global_map_device(ctx, L.PEARSON, L.BF16X3, d_img.data_ptr(), W, H, bb, bb, 0.0, True, d_out.data_ptr())
            torch.cuda.synchronize()
            print(f"# step {i}: contraction {ctx.last_contraction()[0]:.2f} ms", file=sys.stderr, flush=True)
    n = 64 * 1024 * 1024  # 256 MiB of float32
    src = torch.ones(n, device="cuda")
    dst = torch.empty_like(src)
    dst.copy_(src)
    torch.cuda.synchronize()
    print(json.dumps({"profile": True, "copy_bytes_read": 4 * n, "copy_bytes_written": 4 * n}))


def host_cores():
    n = os.environ.get("OMP_NUM_THREADS")
    if n and n.isdigit():
        return int(n)
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:
        return os.cpu_count() or 1


def traffic_from_profile():
    """HBM bytes per launch of the contraction from the committed PMC summary."""
    p = os.path.join(ROOT, "profiles", "pmc_global_pearson_bf16x3_1024_b11.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith("traffic_bytes_per_launch:"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "one GPU only: --gpus must be 1 (row sharding across GPUs is not built)"}))
        sys.exit(2)
    if args.workload == "block":
        block_run(args)
        return
    if args.workload == "adjoint":
        adjoint_run(args)
        return
    if args.workload == "batch":
        batch_run(args)
        return
    if args.workload in ("hu", "phase"):
        pairmaps_run(args)
        return
    if args.workload == "quadtree":
        quadtree_run(args)
        return
    if args.workload == "timelapse":
        timelapse_run(args)
        return
    if args.workload == "tools":
        tools_run(args)
        return
    if args.workload == "vst3d":
        vst3d_run(args)
        return
    if args.workload == "octree":
        octree_run(args)
        return
    if args.workload == "block3d":
        block3d_run(args)
        return
    if args.workload == "vol":
        vol_run(args)
        return
    if args.workload == "vst":
        vst_run(args)
        return
    if args.profile:
        profile_run(args)
        return
    global_run(args)


def pmc_from_profile(key):
    """A figure of the committed PMC summary of the headline contraction."""
    p = os.path.join(ROOT, "profiles", "pmc_global_pearson_bf16x3_1024_b11.txt")
    if not os.path.exists(p):
        return None
    for line in open(p):
        if line.startswith(key + ":"):
            try:
                return float(line.split(":", 1)[1])
            except ValueError:
                return None
    return None


def fp64_errors(O, metric, img, maps, rows, seed):
    """Max/mean abs error of normalised global maps against the FP64 oracle on
    `rows` seeded interior pixels plus each map's 4 smallest and 4 largest
    values: the FP64 accumulators of those pixels only, finalised with the
    relevance mask restricted to them (so the normalisation's min/max come
    from the maps' extremes, which are included)."""
    r = B // 2
    m32, s32 = O.local_stats_f32(img, B, B)
    nv = O.noise_var(img)
    mask, thr, npx = O.relevance_mask(s32, r, r, nv, 0.0)
    idx = np.flatnonzero(mask.ravel() > 0)
    pick = [np.random.default_rng(seed).choice(idx, rows, replace=False)]
    for m in maps.values():
        order = np.argsort(m.ravel()[idx], kind="stable")
        pick += [idx[order[:4]], idx[order[-4:]]]
    px = np.unique(np.concatenate(pick))
    S, Wt = O.global_accum_f64(metric, img, s32, B, B, nv, thr, rows=px)
    sub = np.zeros_like(mask)
    sub.ravel()[px] = 1.0
    ref = O.global_finalize_f64(metric, S, Wt, sub, npx, 0.0, True, B, B).ravel()[px]
    out = {}
    for name, m in maps.items():
        e = np.abs(m.ravel()[px].astype(np.float64) - ref)
        out[name] = {"max": float(e.max()), "mean": float(e.mean())}
    return int(px.size), out


def global_run(args):
    """The headline: global Pearson map 1024x1024, 11x11 (bf16x3 timed), with
    the FP32 form, the all-ordered-pairs split-BF16 form, SSIM and NRMSE in
    both precisions, every map's error against FP64, the roofline of the
    contraction and the CPU baseline."""
    import torch

    import oracle as O
    import sred_amd
    from sred_amd import _lib as L
    from sred_amd import synthetic
    from sred_amd.commands import global_map_device

    img = O.normalize_input(synthetic.lattice_and_blobs(W, seed=2))
    N = (W - 2 * (B // 2)) * (H - 2 * (B // 2))
    Kb = sred_amd.block_size_2d(B, B)
    pairs = float(N) * N

    dev = torch.device("cuda:0")
    stream = torch.cuda.Stream(device=dev)  # default priority
    d_img = torch.from_numpy(img).to(dev)
    d_out = torch.zeros_like(d_img)
    ctx = sred_amd.Context(0)
    ctx.set_stream(stream.cuda_stream)
    maps = {}

    def timed(metric, prec, steps, warmup, symmetric=1):
        ctx.set_option(L.OPT_SYMMETRIC, symmetric)
        step = lambda: global_map_device(ctx, metric, L.BF16X3 if prec == "bf16x3" else L.FP32, d_img.data_ptr(),
                                         W, H, B, B, 0.0, True, d_out.data_ptr())
        try:
            with torch.cuda.stream(stream):
                t, k = _timed(step, steps, warmup, lambda: ctx.last_contraction()[0])
        finally:
            ctx.set_option(L.OPT_SYMMETRIC, 1)
        out = {"ms_per_step": t * 1e3, "contraction_ms": k, "value": pairs / t / 1e9,
               "pairs_evaluated": float(ctx.last_contraction()[2]), "pairs_skipped": float(ctx.last_skipped())}
        maps[(metric, prec, symmetric)] = d_out.cpu().numpy()
        print(f"# metric {metric} {prec}{'' if symmetric else ' all pairs'}: {t * 1e3:.1f} ms/step, "
              f"contraction {k:.1f} ms", file=sys.stderr, flush=True)
        return out

    head = timed(0, "bf16x3", args.steps, args.warmup)
    f32 = timed(0, "fp32", args.fp32_steps, 1)
    full = timed(0, "bf16x3", 1, 0, symmetric=0)
    extra = {}
    for metric, name in ((1, "ssim"), (2, "nrmse")):
        extra[name] = {"bf16x3": timed(metric, "bf16x3", 2, 1), "fp32": timed(metric, "fp32", 1, 0)}
    ctx.close()

    # ---- every map's error against the FP64 oracle (sampled pixels + extremes) ----
    t0 = time.perf_counter()
    errors = {}
    for metric, name in ((0, "pearson"), (1, "ssim"), (2, "nrmse")):
        nsamp, e = fp64_errors(O, metric, img, {p: maps[(metric, p, 1)] for p in ("bf16x3", "fp32")},
                               args.oracle_rows, 123 + metric)
        errors[name] = {"samples": nsamp, **{f"{p}_max": e[p]["max"] for p in e},
                        **{f"{p}_mean": e[p]["mean"] for p in e},
                        "bf16x3_vs_fp32_full_map_max": float(np.abs(maps[(metric, "bf16x3", 1)] -
                                                                    maps[(metric, "fp32", 1)]).max())}
    print(f"# oracle64 errors: {time.perf_counter() - t0:.1f} s", file=sys.stderr, flush=True)
    for name in ("ssim", "nrmse"):
        extra[name]["error_vs_fp64"] = errors[name]
        extra[name]["speedup_vs_fp32"] = extra[name]["fp32"]["ms_per_step"] / extra[name]["bf16x3"]["ms_per_step"]

    # ---- CPU baseline: the oracle's float32 restatement on host cores ----
    cpu = None
    if not args.no_cpu:
        cores = host_cores()
        m32, s32 = O.local_stats_f32(img, B, B)
        nv = O.noise_var(img)
        r = B // 2
        inner = np.zeros((H, W), bool)
        inner[r:H - r, r:W - r] = True
        ref_rows = np.flatnonzero(inner)  # row-major from the first interior pixel
        rows_done, t_cpu, batch = 0, 0.0, 16
        while t_cpu < args.cpu_seconds and rows_done < ref_rows.size:
            rows = ref_rows[rows_done:rows_done + batch]
            t0 = time.perf_counter()
            O.global_accum_f32(O.PEARSON, img, m32, s32, B, B, nv, 0.0, rows=rows, nthreads=cores)
            t_cpu += time.perf_counter() - t0
            rows_done += rows.size
            batch = min(batch * 2, 4096)
        cpu = {"value": rows_done * float(N) / t_cpu / 1e9, "unit": "Gblock-pairs/s", "cores": cores, "kind": "port",
               "sample": f"oracle float32 restatement of kernelGetGlobalPearson2D.cl: first {rows_done} interior "
                         f"reference pixels (row-major) x all {N} test pixels, {t_cpu:.1f} s on {cores} threads"}
        print(f"# cpu: {cpu['value']:.4f} Gpairs/s", file=sys.stderr, flush=True)

    # roofline of the contraction: algorithmic flops N^2 * 2K per launch (every
    # ordered pair, as the reference evaluates them) over the kernel time
    # measured with HIP events on its stream
    k_mfma = head["contraction_ms"]
    achieved = pairs * 2 * Kb / (k_mfma * 1e-3) / 1e12
    kpad = (Kb + 31) // 32 * 32
    issued = head["pairs_evaluated"] * (1.0 - head["pairs_skipped"] / pairs) * 2 * kpad * 3 / (k_mfma * 1e-3) / 1e12
    result = {
        "metric": "Gblock-pairs/s, global Pearson map 1024² b11: MFMA vs FP32 at err≤1e-4",
        "value": head["value"],
        "unit": "Gblock-pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16x3",
        "data": "synthetic: ring lattice (pitch 24, rotated) + 900 Gaussian blobs, noise 0.05, seed 2",
        "config": {"workload": "global repetition, Pearson, 1024x1024, 11x11 blocks (K=81), relevance 0, normalised",
                   "image": [W, H], "block": [B, B], "K": Kb, "pairs": pairs,
                   "precision": "bf16x3 (split BF16 on MFMA, FP32 accumulate, FP64 row sums)"},
        "all_pairs_value": full["value"],
        "fp32": {"value": f32["value"], "ms_per_step": f32["ms_per_step"], "contraction_ms": f32["contraction_ms"]},
        "speedup_vs_fp32": f32["ms_per_step"] / head["ms_per_step"],
        "contraction_ms": k_mfma,
        "pairs_skipped": head["pairs_skipped"], "skip_fraction": head["pairs_skipped"] / pairs,
        "bf16x3_all_pairs": {"value": full["value"], "ms_per_step": full["ms_per_step"],
                             "contraction_ms": full["contraction_ms"],
                             "speedup_vs_fp32": f32["ms_per_step"] / full["ms_per_step"]},
        "error_vs_fp64": errors["pearson"],
        "ssim": extra["ssim"],
        "nrmse": extra["nrmse"],
        "roofline": {"bound": "mfma", "achieved": achieved, "peak": PEAK_BF16_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / PEAK_BF16_TFLOPS, "traffic": traffic_from_profile(),
                     "issued_tflops": issued, "mfma_busy_pmc": pmc_from_profile("mfma_busy_fraction"),
                     "kernel": "k_contract_mfma_sym<Pearson, KS=3>"},
        "cpu_baseline": cpu,
    }
    print(json.dumps(result))


if __name__ == "__main__":
    main()
