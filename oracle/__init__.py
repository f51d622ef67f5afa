"""CPU oracle for SReD's 2D block-correlation path -- TEST INFRASTRUCTURE ONLY.

This package is the checker.  Only ``tests/``, ``__graft_entry__.smoke()`` and
the CPU-baseline leg of ``bench.py`` import it; the product library
(``sred_amd``) never does.

It wraps ``liboracle.so`` (built from ``sred_oracle.c`` by ``oracle/Makefile``)
which restates the reference in two forms:

* ``refsim``   float32-faithful restatement of the OpenCL kernels and Java host
  steps (the reference's own map);
* ``oracle64`` FP64 magnitudes with compensated sums, discrete decisions taken
  from the float32 statistics as the reference takes them.

Every C function cites the reference file:line it follows.
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")
_lib = None

PEARSON, SSIM, NRMSE, ABSDIFF = 0, 1, 2, 3
_F = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_D = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
_I = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_FN3 = ctypes.CFUNCTYPE(ctypes.c_double, ctypes.POINTER(ctypes.c_double), ctypes.c_void_p)


def build():
    """Compile liboracle.so with oracle/Makefile."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        c_int, c_float, c_i64, c_double = ctypes.c_int, ctypes.c_float, ctypes.c_int64, ctypes.c_double
        vp = ctypes.c_void_p
        L.or_footprint.argtypes = [c_int, c_int, vp, vp]
        L.or_java_float_literal.argtypes = [c_float]
        L.or_java_float_literal.restype = c_double
        L.or_normalize_input.argtypes = [_F, c_i64]
        L.or_reference_block.argtypes = [_F, c_int, c_int, _F, ctypes.POINTER(c_float), ctypes.POINTER(c_float)]
        L.or_local_stats_f32.argtypes = [_F, c_int, c_int, c_int, c_int, _F, _F]
        L.or_local_stats_f64.argtypes = [_F, c_int, c_int, c_int, c_int, _D, _D]
        L.or_mean_var_block.argtypes = [_F, c_int, c_int, c_int, c_int, c_int, c_int,
                                        ctypes.POINTER(c_float), ctypes.POINTER(c_float)]
        L.or_noise_var.argtypes = [_F, c_int, c_int]
        L.or_noise_var.restype = c_float
        L.or_relevance_mask.argtypes = [_F, c_int, c_int, c_int, c_int, c_float, c_float, _F,
                                        ctypes.POINTER(c_float), ctypes.POINTER(c_float)]
        L.or_apply_mask.argtypes = [_F, c_i64, _F]
        L.or_normalize_map.argtypes = [_F, c_int, c_int, c_int, c_int, vp]
        L.or_global_accum_f32.argtypes = [c_int, _F, _F, _F, c_int, c_int, c_int, c_int, c_float, c_float,
                                          _F, _F, vp, c_i64, c_int]
        L.or_global_accum_f64.argtypes = [c_int, _F, _F, c_int, c_int, c_int, c_int, c_float, c_float,
                                          _D, _D, vp, c_i64, c_int]
        L.or_global_finalize.argtypes = [c_int, _F, _F, _F, c_float, c_float, c_int, c_int, c_int, c_int, c_int, _F]
        L.or_global_finalize_f64.argtypes = [c_int, _D, _D, _F, c_double, c_float, c_int, c_int, c_int, c_int,
                                             c_int, _D]
        L.or_global_map.argtypes = [c_int, c_int, _F, c_int, c_int, c_int, c_int, c_float, c_int, vp, vp, c_int]
        L.or_block_map.argtypes = [c_int, _F, c_int, c_int, _F, c_int, c_int, c_float, c_float, c_float, c_int, _F]
        L.or_block_raw.argtypes = [c_int, _F, _F, _F, c_int, c_int, _F, c_int, c_int, c_float, c_float, _F]
        L.or_hu1_template.argtypes = [_F, c_int, c_int, c_double]
        L.or_hu1_template.restype = c_double
        L.or_block_hu.argtypes = [_F, c_int, c_int, _F, c_int, c_int, c_double, _F]
        L.or_block_hu.restype = None
        L.or_hu_prepare_template.argtypes = [_F, c_int, c_int, c_double, _F]
        L.or_hu_prepare_template.restype = None
        L.or_hu_invariant.argtypes = [_F, c_int, c_int, c_int, c_int, c_int, c_float, c_double]
        L.or_hu_invariant.restype = c_float
        L.or_hu_invariants.argtypes = [_F, _F, c_int, c_int, c_int, c_int, c_double, _F]
        L.or_hu_invariants.restype = None
        L.or_hu_weight.argtypes = [c_float, c_float, c_double]
        L.or_hu_weight.restype = c_double
        L.or_global_hu_sums.argtypes = [_F, _F, c_int, c_int, c_int, c_int, c_double, vp, c_i64, _D, c_int]
        L.or_global_hu_sums.restype = None
        L.or_phase_peak_blocks.argtypes = [_D, _D, c_int, c_int, c_double, vp, ctypes.POINTER(c_int),
                                           ctypes.POINTER(c_int)]
        L.or_phase_peak_blocks.restype = c_int
        L.or_phase_weight.argtypes = [c_float, c_float]
        L.or_phase_weight.restype = c_double
        L.or_phase_map.argtypes = [_F, _F, _F, c_int, c_int, c_int, c_int, c_double, vp, c_i64, _D, c_int]
        L.or_phase_map.restype = None
        L.or_gat.argtypes = [_F, c_i64, c_double, c_double, c_double]
        L.or_gat.restype = None
        L.or_gat_objective.argtypes = [_F, c_int, c_int, c_double, c_double, c_double]
        L.or_gat_objective.restype = c_double
        L.or_nelder_mead3.argtypes = [_FN3, vp, _D, _D, c_int, _D]
        L.or_nelder_mead3.restype = c_double
        L.or_gat_minimize.argtypes = [_F, c_int, c_int, c_double, c_double, c_double, c_int, _D]
        L.or_gat_minimize.restype = c_double
        L.or_beta_reg.argtypes = [c_double, c_double, c_double]
        L.or_beta_reg.restype = c_double
        L.or_f_quantile.argtypes = [c_double, c_double, c_double]
        L.or_f_quantile.restype = c_double
        L.or_simple_regression.argtypes = [_D, _D, c_i64, ctypes.POINTER(c_double), ctypes.POINTER(c_double)]
        L.or_simple_regression.restype = c_int
        L.or_qt_criterion.argtypes = [_F, c_int, c_int, c_int, c_int, c_int, c_int, ctypes.POINTER(c_float),
                                      ctypes.POINTER(c_float)]
        L.or_qt_criterion.restype = None
        L.or_qt_robust_mean.argtypes = [_F, c_int, c_int, c_int, c_int, c_int, c_int]
        L.or_qt_robust_mean.restype = c_float
        L.or_qt_lts_variance.argtypes = [_F, c_int, c_int, c_int, c_int, c_int, c_float, c_float]
        L.or_qt_lts_variance.restype = c_float
        L.or_quadtree_gat.argtypes = [_F, c_int, c_int, c_int, c_float, c_int, c_float, c_int, vp, _F, _F,
                                      ctypes.POINTER(c_int), ctypes.POINTER(c_double), ctypes.POINTER(c_double)]
        L.or_quadtree_gat.restype = c_int
        L.or_gat_tree.argtypes = [_F, c_i64, c_double, c_double]
        L.or_gat_tree.restype = None
        L.or_oc_criterion.argtypes = [_F, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int,
                                      ctypes.POINTER(c_float), ctypes.POINTER(c_float)]
        L.or_oc_criterion.restype = None
        L.or_octree_gat.argtypes = [_F, c_int, c_int, c_int, c_int, c_float, c_int, c_float, c_int, vp, _F, _F,
                                    ctypes.POINTER(c_int), ctypes.POINTER(c_double), ctypes.POINTER(c_double)]
        L.or_octree_gat.restype = c_int
        L.or_gat_tree3d.argtypes = [_F, c_i64, c_float, c_float]
        L.or_gat_tree3d.restype = None
        L.or_relevance_mask_stack.argtypes = [_F, c_int, c_int, c_int, c_int, c_float, c_float, c_float, c_int, _F,
                                              _F, ctypes.POINTER(c_int)]
        L.or_relevance_mask_stack.restype = c_int
        L.or_gat_blocks3d.argtypes = [c_int, c_int, c_int] + [ctypes.POINTER(c_int)] * 6
        L.or_gat_blocks3d.restype = c_int
        L.or_gat_objective3d.argtypes = [_F, c_int, c_int, c_int, c_double, c_double, c_double]
        L.or_gat_objective3d.restype = c_double
        L.or_gat_minimize3d.argtypes = [_F, c_int, c_int, c_int, c_double, c_double, c_double, c_int, _D]
        L.or_gat_minimize3d.restype = c_double
        L.or_oc_robust_mean.argtypes = [_F, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int]
        L.or_oc_robust_mean.restype = c_float
        L.or_oc_lts_variance.argtypes = [_F, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_int, c_float,
                                         c_float]
        L.or_oc_lts_variance.restype = c_float
        L.or_tool_prep.argtypes = [_F, c_int, c_int, c_int, _D]
        L.or_tool_prep.restype = None
        L.or_noise_var_double.argtypes = [_F, c_int, c_int]
        L.or_noise_var_double.restype = c_float
        L.or_relevance_mask_tool.argtypes = [_F, c_int, c_int, c_int, c_int, c_float, _F, ctypes.POINTER(c_float),
                                             ctypes.POINTER(c_float)]
        L.or_relevance_mask_tool.restype = c_int
        L.or_local_var_f32.argtypes = [_F, c_int, c_int, c_int, c_int, _F]
        L.or_local_var_f32.restype = None
        L.or_optimise_relevance_mask.argtypes = [_F, c_int, c_int, c_int, c_int, c_float, c_int,
                                                 ctypes.POINTER(c_float), ctypes.POINTER(c_float)]
        L.or_optimise_relevance_mask.restype = c_int
        L.or_footprint3d.argtypes = [c_int, c_int, c_int, vp, vp, vp]
        L.or_footprint3d.restype = c_int
        L.or_local_stats3d_f32.argtypes = [_F, c_int, c_int, c_int, c_int, c_int, c_int, _F, _F]
        L.or_local_stats3d_f32.restype = None
        L.or_vol_noise_var.argtypes = [_F, c_int, c_int, c_int, c_int, c_int, c_int]
        L.or_vol_noise_var.restype = c_float
        L.or_global_pearson3d.argtypes = [_F, c_int, c_int, c_int, c_int, c_int, c_int, c_float, c_int, c_int, _D,
                                          vp, vp, vp, ctypes.POINTER(c_float), ctypes.POINTER(c_float)]
        L.or_global_pearson3d.restype = c_int
        L.or_global_repetition3d.argtypes = [c_int, _F, c_int, c_int, c_int, c_int, c_int, c_int, c_float, c_int,
                                             c_int, _D, vp, vp, vp, ctypes.POINTER(c_float), ctypes.POINTER(c_float)]
        L.or_global_repetition3d.restype = c_int
        L.or_reference_block3d.argtypes = [_F, c_int, c_int, c_int, _F, ctypes.POINTER(c_float),
                                           ctypes.POINTER(c_float)]
        L.or_reference_block3d.restype = c_int
        L.or_block_raw3d.argtypes = [c_int, _F, _F, _F, c_int, c_int, c_int, _F, c_int, c_int, c_int, c_float,
                                     c_float, _F]
        L.or_block_raw3d.restype = c_int
        L.or_block_map3d.argtypes = [c_int, _F, c_int, c_int, c_int, _F, c_int, c_int, c_int, c_float, c_float,
                                     c_float, c_int, _F]
        L.or_block_map3d.restype = c_int
        _lib = L
    return _lib


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def footprint(bw, bh):
    """(K, dx, dy) of the elliptical block, reference row-major order."""
    k = lib().or_footprint(bw, bh, None, None)
    dx = (ctypes.c_int * k)()
    dy = (ctypes.c_int * k)()
    lib().or_footprint(bw, bh, dx, dy)
    return k, np.array(dx[:], dtype=np.int32), np.array(dy[:], dtype=np.int32)


OR_MAX_K = 3200  # largest footprint the C restatement holds per thread (sred_oracle.h)


def _check_k(bw, bh):
    k = lib().or_footprint(bw, bh, None, None)
    if k > OR_MAX_K:
        raise ValueError(f"block {bw}x{bh} has {k} pixels, more than the oracle's {OR_MAX_K}")


def java_float_literal(f):
    return lib().or_java_float_literal(float(np.float32(f)))


def normalize_input(img):
    out = _f32(img).copy()
    lib().or_normalize_input(out.reshape(-1), out.size)
    return out


def reference_block(block):
    block = _f32(block)
    bh, bw = block.shape
    px = np.zeros(bw * bh, np.float32)
    m, s = ctypes.c_float(), ctypes.c_float()
    k = lib().or_reference_block(block.reshape(-1), bw, bh, px, ctypes.byref(m), ctypes.byref(s))
    if k < 0:
        return None
    return px[:k].copy(), m.value, s.value


def local_stats_f32(img, bw, bh):
    _check_k(bw, bh)
    img = _f32(img)
    h, w = img.shape
    m = np.zeros_like(img)
    s = np.zeros_like(img)
    lib().or_local_stats_f32(img.reshape(-1), w, h, bw, bh, m.reshape(-1), s.reshape(-1))
    return m, s


def local_stats_f64(img, bw, bh):
    img = _f32(img)
    h, w = img.shape
    m = np.zeros(img.shape, np.float64)
    v = np.zeros(img.shape, np.float64)
    lib().or_local_stats_f64(img.reshape(-1), w, h, bw, bh, m.reshape(-1), v.reshape(-1))
    return m, v


def mean_var_block(img, x0, y0, x1, y1):
    img = _f32(img)
    h, w = img.shape
    m, v = ctypes.c_float(), ctypes.c_float()
    rc = lib().or_mean_var_block(img.reshape(-1), w, h, x0, y0, x1, y1, ctypes.byref(m), ctypes.byref(v))
    if rc:
        raise ValueError("invalid block")
    return m.value, v.value


def noise_var(img):
    img = _f32(img)
    h, w = img.shape
    return lib().or_noise_var(img.reshape(-1), w, h)


def relevance_mask(stds, brw, brh, noise_variance, c):
    stds = _f32(stds)
    h, w = stds.shape
    mask = np.zeros_like(stds)
    thr, npx = ctypes.c_float(), ctypes.c_float()
    lib().or_relevance_mask(stds.reshape(-1), w, h, brw, brh, noise_variance, c, mask.reshape(-1),
                            ctypes.byref(thr), ctypes.byref(npx))
    return mask, thr.value, npx.value


def apply_mask(img, mask):
    out = _f32(img).copy()
    lib().or_apply_mask(out.reshape(-1), out.size, _f32(mask).reshape(-1))
    return out


def normalize_map(img, brw, brh, mask=None):
    out = _f32(img).copy()
    h, w = out.shape
    mk = None if mask is None else _f32(mask)
    rc = lib().or_normalize_map(out.reshape(-1), w, h, brw, brh,
                                None if mk is None else mk.ctypes.data_as(ctypes.c_void_p))
    if rc:
        raise ValueError("Relevance mask must be binary (0.0f or 1.0f)")
    return out


def _rows(rows):
    if rows is None:
        return None, 0, None
    r = np.ascontiguousarray(rows, dtype=np.int32)
    return r.ctypes.data_as(ctypes.c_void_p), r.size, r


def global_accum_f32(metric, img, means, stds, bw, bh, filter_param, threshold, rows=None, nthreads=0):
    _check_k(bw, bh)
    img = _f32(img)
    h, w = img.shape
    S = np.zeros_like(img)
    W = np.zeros_like(img)
    rp, nr, _keep = _rows(rows)
    lib().or_global_accum_f32(metric, img.reshape(-1), _f32(means).reshape(-1), _f32(stds).reshape(-1), w, h,
                              bw, bh, filter_param, threshold, S.reshape(-1), W.reshape(-1), rp, nr, nthreads)
    return S, W


def global_accum_f64(metric, img, stds32, bw, bh, filter_param, threshold, rows=None, nthreads=0):
    _check_k(bw, bh)
    img = _f32(img)
    h, w = img.shape
    S = np.zeros(img.shape, np.float64)
    W = np.zeros(img.shape, np.float64)
    rp, nr, _keep = _rows(rows)
    lib().or_global_accum_f64(metric, img.reshape(-1), _f32(stds32).reshape(-1), w, h, bw, bh, filter_param,
                              threshold, S.reshape(-1), W.reshape(-1), rp, nr, nthreads)
    return S, W


def global_finalize(metric, S, W, mask, n_pixels, c, normalize, bw, bh):
    S = _f32(S)
    h, w = S.shape
    out = np.zeros_like(S)
    rc = lib().or_global_finalize(metric, S.reshape(-1), _f32(W).reshape(-1), _f32(mask).reshape(-1), n_pixels,
                                  c, int(normalize), w, h, bw, bh, out.reshape(-1))
    if rc:
        raise ValueError("finalize failed")
    return out


def global_finalize_f64(metric, S, W, mask, n_pixels, c, normalize, bw, bh):
    S = np.ascontiguousarray(S, np.float64)
    h, w = S.shape
    out = np.zeros_like(S)
    rc = lib().or_global_finalize_f64(metric, S.reshape(-1), np.ascontiguousarray(W, np.float64).reshape(-1),
                                      _f32(mask).reshape(-1), float(n_pixels), c, int(normalize), w, h, bw, bh,
                                      out.reshape(-1))
    if rc:
        raise ValueError("finalize failed")
    return out


def global_map(metric, img, bw, bh, c=0.0, normalize=True, fp64=False, nthreads=0):
    """Full global-repetition map (refsim float32, or oracle64 with fp64=True)."""
    _check_k(bw, bh)
    img = _f32(img)
    h, w = img.shape
    if fp64:
        out = np.zeros(img.shape, np.float64)
        rc = lib().or_global_map(metric, 1, img.reshape(-1), w, h, bw, bh, c, int(normalize), None,
                                 out.ctypes.data_as(ctypes.c_void_p), nthreads)
    else:
        out = np.zeros(img.shape, np.float32)
        rc = lib().or_global_map(metric, 0, img.reshape(-1), w, h, bw, bh, c, int(normalize),
                                 out.ctypes.data_as(ctypes.c_void_p), None, nthreads)
    if rc:
        raise ValueError("global map failed")
    return out


def block_raw(metric, img, means, stds, tmpl, bw, bh, tmpl_mean, tmpl_std):
    img = _f32(img)
    h, w = img.shape
    out = np.zeros_like(img)
    rc = lib().or_block_raw(metric, img.reshape(-1), _f32(means).reshape(-1), _f32(stds).reshape(-1), w, h,
                            _f32(tmpl), bw, bh, tmpl_mean, tmpl_std, out.reshape(-1))
    if rc:
        raise ValueError("block raw failed")
    return out


def block_map(metric, img, tmpl, bw, bh, tmpl_mean, tmpl_std, c=0.0, normalize=True):
    img = _f32(img)
    h, w = img.shape
    out = np.zeros_like(img)
    rc = lib().or_block_map(metric, img.reshape(-1), w, h, _f32(tmpl), bw, bh, tmpl_mean, tmpl_std, c,
                            int(normalize), out.reshape(-1))
    if rc:
        raise ValueError("block map failed")
    return out


def hu_eps():
    """The double the Hu kernel's EPSILON macro denotes (Java ""+1e-7f)."""
    return java_float_literal(np.float32(1e-7))


def hu1_template(tmpl, eps=None):
    """hu1 of a template rectangle (bh, bw), kernelGetPatchHu2D.cl:36-101."""
    t = _f32(tmpl)
    bh, bw = t.shape
    return lib().or_hu1_template(t.reshape(-1), bw, bh, hu_eps() if eps is None else eps)


def block_hu(img, tmpl, eps=None):
    """Hu template map, kernelGetPatchHu2D.cl (border 0)."""
    img = _f32(img)
    t = _f32(tmpl)
    h, w = img.shape
    bh, bw = t.shape
    out = np.zeros_like(img)
    lib().or_block_hu(img.reshape(-1), w, h, t.reshape(-1), bw, bh, hu_eps() if eps is None else eps,
                      out.reshape(-1))
    return out


def hu_prepare_template(blk, eps=None):
    """The build's Hu template preparation (window steps on the template)."""
    b = _f32(blk)
    bh, bw = b.shape
    out = np.zeros_like(b)
    lib().or_hu_prepare_template(b.reshape(-1), bw, bh, hu_eps() if eps is None else eps, out.reshape(-1))
    return out


def hu_invariants(img, means, bw, bh, eps=None):
    """hu = inv20 + inv02 of kernelGetHuMap.cl:146-172 at every interior pixel (border 0)."""
    img, means = _f32(img), _f32(means)
    h, w = img.shape
    out = np.zeros_like(img)
    lib().or_hu_invariants(img.reshape(-1), means.reshape(-1), w, h, bw, bh, hu_eps() if eps is None else eps,
                           out.reshape(-1))
    return out


def hu_weight(a, b, eps=None):
    """getExpDecayWeight of kernelGetHuMap.cl:136-144 in FP64."""
    return lib().or_hu_weight(float(a), float(b), hu_eps() if eps is None else eps)


def global_hu_sums(hu, stds, bw, bh, rows=None, eps=None, nthreads=0):
    """FP64 global Hu map sums (kernelGetHuMap.cl:87-133) of given invariants."""
    hu, stds = _f32(hu), _f32(stds)
    h, w = hu.shape
    out = np.zeros((h, w), np.float64)
    rp, nr, _keep = _rows(rows)
    lib().or_global_hu_sums(hu.reshape(-1), stds.reshape(-1), w, h, bw, bh, hu_eps() if eps is None else eps,
                            rp, nr, out.reshape(-1), nthreads)
    return out


def global_hu_map(img, means, stds, bw, bh, rows=None, eps=None, nthreads=0):
    """Global Hu map (kernelGetHuMap.cl) from local statistics, FP64 sums."""
    return global_hu_sums(hu_invariants(img, means, bw, bh, eps), stds, bw, bh, rows, eps, nthreads)


def phase_peak_blocks(bi, bj, eps=None):
    """r = Re IDFT(normalised cross power) of two normalised blocks (bh, bw) and
    the wrapped peak (dx, dy); found is False when no value of r is positive."""
    bi = np.ascontiguousarray(bi, np.float64)
    bj = np.ascontiguousarray(bj, np.float64)
    bh, bw = bi.shape
    r = np.zeros((bh, bw), np.float64)
    dx, dy = ctypes.c_int(), ctypes.c_int()
    f = lib().or_phase_peak_blocks(bi.reshape(-1), bj.reshape(-1), bw, bh, hu_eps() if eps is None else eps,
                                   r.ctypes.data, ctypes.byref(dx), ctypes.byref(dy))
    return r, (dx.value, dy.value), bool(f)


def phase_weight(a, b):
    """getExpDecayWeight of kernelGetPhaseCorrelationMap.cl:167-178 (float)."""
    return lib().or_phase_weight(float(a), float(b))


def phase_map(img, means, stds, bw, bh, rows=None, eps=None, nthreads=0):
    """Phase correlation map (textbook specification of
    kernelGetPhaseCorrelationMap.cl), FP64."""
    img, means, stds = _f32(img), _f32(means), _f32(stds)
    h, w = img.shape
    out = np.zeros((h, w), np.float64)
    rp, nr, _keep = _rows(rows)
    lib().or_phase_map(img.reshape(-1), means.reshape(-1), stds.reshape(-1), w, h, bw, bh,
                       hu_eps() if eps is None else eps, rp, nr, out.reshape(-1), nthreads)
    return out


def gat(img, gain, sigma, offset):
    """VarianceStabilisingTransform2D_.getGAT (a copy)."""
    out = _f32(img).copy()
    lib().or_gat(out.reshape(-1), out.size, float(gain), float(sigma), float(offset))
    return out


def gat_objective(img, gain, sigma, offset):
    """GATMinimizer2D.userFunction (fixed 64-lane summation order)."""
    img = _f32(img)
    h, w = img.shape
    return lib().or_gat_objective(img.reshape(-1), w, h, float(gain), float(sigma), float(offset))


def nelder_mead3(f, x0, d, max_iter):
    """The oracle's Nelder-Mead on a Python function of 3 parameters."""
    cb = _FN3(lambda x, _c: float(f(np.array([x[0], x[1], x[2]]))))
    x = np.zeros(3)
    fx = lib().or_nelder_mead3(cb, None, np.asarray(x0, np.float64), np.asarray(d, np.float64), int(max_iter), x)
    return x, fx


def gat_minimize(img, gain0=1.0, sigma0=10.0, offset0=100.0, max_iter=5000):
    """GATMinimizer2D.run with the oracle's simplex: (gain, sigma, offset), objective."""
    img = _f32(img)
    h, w = img.shape
    x = np.zeros(3)
    fx = lib().or_gat_minimize(img.reshape(-1), w, h, gain0, sigma0, offset0, int(max_iter), x)
    return x, fx


def beta_reg(a, b, x):
    """Regularised incomplete beta I_x(a, b)."""
    return lib().or_beta_reg(float(a), float(b), float(x))


def f_quantile(d1, d2, p):
    """Inverse CDF of the F(d1, d2) distribution (FDistribution.inverseCumulativeProbability)."""
    return lib().or_f_quantile(float(d1), float(d2), float(p))


def simple_regression(x, y):
    """commons-math3 SimpleRegression with intercept: (slope, intercept)."""
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    s, i = ctypes.c_double(), ctypes.c_double()
    lib().or_simple_regression(x, y, x.size, ctypes.byref(s), ctypes.byref(i))
    return s.value, i.value


def qt_criterion(img, x0, y0, nw, nh):
    """QuadTree_.computeCriterion's two variances: (block variance, noise variance)."""
    img = _f32(img)
    h, w = img.shape
    bv, nv = ctypes.c_float(), ctypes.c_float()
    lib().or_qt_criterion(img.reshape(-1), w, h, x0, y0, nw, nh, ctypes.byref(bv), ctypes.byref(nv))
    return bv.value, nv.value


def qt_robust_mean(img, x0, y0, nw, nh, max_iter=50):
    img = _f32(img)
    return lib().or_qt_robust_mean(img.reshape(-1), img.shape[1], x0, y0, nw, nh, int(max_iter))


def qt_lts_variance(img, x0, y0, nw, nh, mean, frac=0.75):
    img = _f32(img)
    return lib().or_qt_lts_variance(img.reshape(-1), img.shape[1], x0, y0, nw, nh, float(mean), float(frac))


class QuadTreeResult:
    """Leaves in pre-order (x, y, w, h), their robust means and LTS variances, g0 and eDC."""

    def __init__(self, rects, means, variances, g0, edc):
        self.rects, self.means, self.variances, self.g0, self.edc = rects, means, variances, g0, edc


def quadtree_gat(img, min_size=4, alpha=0.01, max_iter=50, frac=0.75):
    """Utils.getInputImage2D's Quad/Octree estimate; raises ValueError(status) on the reference's errors."""
    img = _f32(img)
    h, w = img.shape
    cap = max(16, (w * h) // 2)
    rects = np.zeros((cap, 4), np.int32)
    means = np.zeros(cap, np.float32)
    var = np.zeros(cap, np.float32)
    n = ctypes.c_int()
    g0, edc = ctypes.c_double(), ctypes.c_double()
    rc = lib().or_quadtree_gat(img.reshape(-1), w, h, int(min_size), float(alpha), int(max_iter), float(frac), cap,
                               rects.ctypes.data, means, var, ctypes.byref(n), ctypes.byref(g0), ctypes.byref(edc))
    if rc != 0:
        raise ValueError(rc)
    k = n.value
    return QuadTreeResult(rects[:k].copy(), means[:k].copy(), var[:k].copy(), g0.value, edc.value)


def gat_tree(img, g0, edc):
    """QuadTree_.applyGATtree (a copy)."""
    out = _f32(img).copy()
    lib().or_gat_tree(out.reshape(-1), out.size, float(g0), float(edc))
    return out


def block_timelapse(metric, stack, tmpl_px, bw, bh, tmpl_mean, tmpl_std, c, normalize):
    """The time-lapse branch of BlockRepetition2D_ (BlockRepetition2D_.java:95-137)
    without stabilisation, composed of the oracle's steps: the stack
    normalised as a whole (Utils.normalizeArray), each frame normalised
    (getInputImage2D(float[])) and mapped (block_map), the maps normalised as
    a whole when normalize."""
    st = _f32(stack)
    d, h, w = st.shape
    if d < 3:
        raise ValueError("Image must have at least 3 slices.")
    flat = normalize_input(st.reshape(-1)).reshape(d, h, w)
    out = np.zeros_like(flat)
    for t in range(d):
        out[t] = block_map(metric, normalize_input(flat[t]), tmpl_px, bw, bh, tmpl_mean, tmpl_std, c, normalize)
    if normalize:
        out = normalize_input(out.reshape(-1)).reshape(d, h, w)
    return out


def tool_prep(img, max_iter=5000):
    """The relevance-mask tools' input step: (prepared image, GAT fit)."""
    out = _f32(img).copy()
    h, w = out.shape
    p = np.zeros(3)
    lib().or_tool_prep(out.reshape(-1), w, h, int(max_iter), p)
    return out, p


def noise_var_double(img):
    img = _f32(img)
    h, w = img.shape
    return lib().or_noise_var_double(img.reshape(-1), w, h)


def relevance_mask_tool(img, bw, bh, c):
    """RelevanceMask2D_ after its input step: (mask, noise variance, threshold)."""
    img = _f32(img)
    h, w = img.shape
    mask = np.zeros_like(img)
    nv, thr = ctypes.c_float(), ctypes.c_float()
    if lib().or_relevance_mask_tool(img.reshape(-1), w, h, bw, bh, float(c), mask.reshape(-1), ctypes.byref(nv),
                                    ctypes.byref(thr)) != 0:
        raise ValueError("blocks must be odd and square")
    return mask, nv.value, thr.value


def local_var_f32(img, bw, bh):
    img = _f32(img)
    h, w = img.shape
    out = np.zeros_like(img)
    lib().or_local_var_f32(img.reshape(-1), w, h, bw, bh, out.reshape(-1))
    return out


def optimise_relevance_mask(img, bw, bh, step, max_steps=100000):
    """OptimiseRelevanceMask_ after its input step: (all-ones constant, all-zeros constant)."""
    img = _f32(img)
    h, w = img.shape
    a, b = ctypes.c_float(), ctypes.c_float()
    rc = lib().or_optimise_relevance_mask(img.reshape(-1), w, h, bw, bh, float(step), int(max_steps), ctypes.byref(a),
                                          ctypes.byref(b))
    if rc != 0:
        raise ValueError(rc)
    return a.value, b.value


def footprint3d(bw, bh, bd):
    """(K, dx, dy, dz) of the ellipsoid block, z-y-x order."""
    k = lib().or_footprint3d(bw, bh, bd, None, None, None)
    d = np.zeros((3, k), np.int32)
    lib().or_footprint3d(bw, bh, bd, d[0].ctypes.data, d[1].ctypes.data, d[2].ctypes.data)
    return k, d[0], d[1], d[2]


def local_stats3d_f32(vol, bw, bh, bd):
    vol = _f32(vol)
    d, h, w = vol.shape
    m = np.zeros_like(vol)
    s = np.zeros_like(vol)
    lib().or_local_stats3d_f32(vol.reshape(-1), w, h, d, bw, bh, bd, m.reshape(-1), s.reshape(-1))
    return m, s


def vol_noise_var(stds, bw, bh, bd):
    stds = _f32(stds)
    d, h, w = stds.shape
    return lib().or_vol_noise_var(stds.reshape(-1), w, h, d, bw // 2, bh // 2, bd // 2)


class Repetition3D:
    """A 3D global map and its intermediates (float form)."""

    def __init__(self, out, S, W, mask, noise_var, n_pixels):
        self.out, self.S, self.W, self.mask, self.noise_var, self.n_pixels = out, S, W, mask, noise_var, n_pixels


Pearson3D = Repetition3D


def global_repetition3d(vol, metric, bw, bh, bd, c=0.0, normalize=True, fp64=False):
    """GlobalRepetition3D_ with `metric` (PEARSON, SSIM, NRMSE, ABSDIFF) as
    intended: the kernels' float form, or FP64 with fp64=True."""
    vol = _f32(vol)
    d, h, w = vol.shape
    out = np.zeros(vol.shape, np.float64)
    S = np.zeros_like(vol)
    W = np.zeros_like(vol)
    mask = np.zeros_like(vol)
    nv, npx = ctypes.c_float(), ctypes.c_float()
    rc = lib().or_global_repetition3d(int(metric), vol.reshape(-1), w, h, d, bw, bh, bd, float(c), int(normalize),
                                      int(fp64), out.reshape(-1), S.ctypes.data, W.ctypes.data, mask.ctypes.data,
                                      ctypes.byref(nv), ctypes.byref(npx))
    if rc != 0:
        raise ValueError(f"unknown metric {metric}")
    return Repetition3D(out, S, W, mask, nv.value, npx.value)


def global_pearson3d(vol, bw, bh, bd, c=0.0, normalize=True, fp64=False):
    """GlobalRepetition3D_ with Pearson's R as intended (refsim float, or FP64 with fp64=True)."""
    return global_repetition3d(vol, PEARSON, bw, bh, bd, c, normalize, fp64)


def reference_block3d(block):
    """Utils.getReferenceBlock3D on a (depth, H, W) block: (pixels, mean, std)
    with the pixels in the ellipsoid's z, y, x order; None for even sides or
    a depth below 3."""
    blk = _f32(block)
    bd, bh, bw = blk.shape
    px = np.zeros(bw * bh * bd, np.float32)
    m, sd = ctypes.c_float(), ctypes.c_float()
    k = lib().or_reference_block3d(blk.reshape(-1), bw, bh, bd, px, ctypes.byref(m), ctypes.byref(sd))
    if k < 0:
        return None
    return px[:k].copy(), m.value, sd.value


def block_raw3d(metric, vol, means, stds, tmpl, bw, bh, bd, tmpl_mean, tmpl_std):
    """Raw output of the 3D block kernel of `metric` on a (depth, H, W) volume."""
    vol = _f32(vol)
    d, h, w = vol.shape
    out = np.zeros_like(vol)
    rc = lib().or_block_raw3d(int(metric), vol.reshape(-1), _f32(means).reshape(-1), _f32(stds).reshape(-1), w, h,
                              d, _f32(tmpl), bw, bh, bd, float(tmpl_mean), float(tmpl_std), out.reshape(-1))
    if rc != 0:
        raise ValueError("unknown metric or block above 8192 voxels")
    return out


def block_map3d(metric, vol, tmpl, bw, bh, bd, tmpl_mean, tmpl_std, c=0.0, normalize=True):
    """CLUtils.calculateBlockRepetitionMap3D as intended on a normalised
    (depth, H, W) volume (float form)."""
    vol = _f32(vol)
    d, h, w = vol.shape
    out = np.zeros_like(vol)
    rc = lib().or_block_map3d(int(metric), vol.reshape(-1), w, h, d, _f32(tmpl), bw, bh, bd, float(tmpl_mean),
                              float(tmpl_std), float(c), int(normalize), out.reshape(-1))
    if rc != 0:
        raise ValueError("unknown metric or block above 8192 voxels")
    return out


def oc_criterion(vol, x0, y0, z0, nw, nh, nd):
    """OcTree_.computeCriterion's two variances: (block variance, noise variance)."""
    vol = _f32(vol)
    d, h, w = vol.shape
    bv, nv = ctypes.c_float(), ctypes.c_float()
    lib().or_oc_criterion(vol.reshape(-1), w, h, d, x0, y0, z0, nw, nh, nd, ctypes.byref(bv), ctypes.byref(nv))
    return bv.value, nv.value


class OcTreeResult:
    """Leaves in pre-order (x, y, z, w, h, d), their robust means and LTS variances, g0 and eDC (float values)."""

    def __init__(self, boxes, means, variances, g0, edc):
        self.boxes, self.means, self.variances, self.g0, self.edc = boxes, means, variances, g0, edc


def octree_gat(vol, min_size=4, alpha=0.01, max_iter=50, frac=0.75):
    """The OcTree_ estimate of a (depth, H, W) volume; raises ValueError(status) on the reference's errors."""
    vol = _f32(vol)
    d, h, w = vol.shape
    cap = max(16, vol.size // 2)
    boxes = np.zeros((cap, 6), np.int32)
    means = np.zeros(cap, np.float32)
    var = np.zeros(cap, np.float32)
    n, g0, edc = ctypes.c_int(), ctypes.c_double(), ctypes.c_double()
    rc = lib().or_octree_gat(vol.reshape(-1), w, h, d, int(min_size), float(alpha), int(max_iter), float(frac), cap,
                             boxes.ctypes.data, means, var, ctypes.byref(n), ctypes.byref(g0), ctypes.byref(edc))
    if rc != 0:
        raise ValueError(rc)
    k = n.value
    return OcTreeResult(boxes[:k].copy(), means[:k].copy(), var[:k].copy(), g0.value, edc.value)


def gat_tree3d(vol, g0, edc):
    """OcTree_.applyGATtree with float g0, eDC (a copy)."""
    out = _f32(vol).copy()
    lib().or_gat_tree3d(out.reshape(-1), out.size, float(g0), float(edc))
    return out


def gat_blocks3d(w, h, d):
    """GATMinimizer3D's block grid: (bw, bh, bd, nbx, nby, nbz)."""
    v = [ctypes.c_int() for _ in range(6)]
    if lib().or_gat_blocks3d(w, h, d, *[ctypes.byref(x) for x in v]):
        raise ValueError("the reference's block search does not terminate for this volume")
    return tuple(x.value for x in v)


def gat_objective3d(vol, gain, sigma, offset):
    """GATMinimizer3D.userFunction (device sum order)."""
    vol = _f32(vol)
    d, h, w = vol.shape
    return lib().or_gat_objective3d(vol.reshape(-1), w, h, d, float(gain), float(sigma), float(offset))


def gat_minimize3d(vol, gain0=1.0, sigma0=100.0, offset0=10.0, max_iter=5000):
    """The simplex over the 3D objective: ((gain, sigma, offset), error)."""
    vol = _f32(vol)
    d, h, w = vol.shape
    gat_blocks3d(w, h, d)  # raises where the reference never ends
    p = np.zeros(3)
    err = lib().or_gat_minimize3d(vol.reshape(-1), w, h, d, float(gain0), float(sigma0), float(offset0),
                                  int(max_iter), p)
    return tuple(p), err


def relevance_mask_stack(img, bw, bh, lower, upper, step, cap=4096):
    """RelevanceMaskStack2D_ after the input step: (masks (n, h, w), constants (n,))."""
    img = _f32(img)
    h, w = img.shape
    n = ctypes.c_int()
    rc = lib().or_relevance_mask_stack(img.reshape(-1), w, h, bw, bh, float(lower), float(upper), float(step), 0,
                                       np.zeros(1, np.float32), np.zeros(1, np.float32), ctypes.byref(n))
    if rc == -1:
        raise ValueError("invalid block or constants")
    k = n.value
    masks = np.zeros((k, h, w), np.float32)
    consts = np.zeros(k, np.float32)
    rc = lib().or_relevance_mask_stack(img.reshape(-1), w, h, bw, bh, float(lower), float(upper), float(step), k,
                                       masks.reshape(-1), consts, ctypes.byref(n))
    if rc != 0:
        raise ValueError(rc)
    return masks, consts
