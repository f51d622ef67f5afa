"""Mirrors of SReD's relevance-mask tools (RelevanceMask2D_,
RelevanceMaskStack2D_, OptimiseRelevanceMask_, CalculateLocalMeans_) over
the C ABI.

Each plugin first stabilises and normalises the image the same way
(GATMinimizer2D from (0, 100, 0), getGAT with the fitted doubles, a double
min-max normalisation); prepare_input does that step, and the functions
below take the raw image and run it first.
"""
import ctypes

import numpy as np

from . import _lib as L
from . import clutils, utils


def _ctx(resources):
    return (L.Context(0), True) if resources is None else (resources, False)


def prepare_input(image, max_iter=5000, resources=None):
    """The tools' input step: (prepared float32 image, (gain, sigma, offset))."""
    a = L.f32(np.array(image, dtype=np.float32, copy=True))
    h, w = a.shape
    p = (ctypes.c_double * 3)()
    res, own = _ctx(resources)
    try:
        L.check(L.load().sred_tool_prep_2d(res.handle, L.fptr(a), w, h, int(max_iter), p))
    finally:
        if own:
            res.close()
    return a, (p[0], p[1], p[2])


def relevance_mask_2d(image, block_width=3, block_height=3, filter_constant=0.0, max_iter=5000, resources=None):
    """RelevanceMask2D_.run (RelevanceMask2D_.java:30-250): the relevance mask
    of the stabilised image as (mask, noise variance, threshold)."""
    res, own = _ctx(resources)
    try:
        a, _ = prepare_input(image, max_iter, res)
        h, w = a.shape
        mask = np.zeros_like(a)
        nv, thr = ctypes.c_float(), ctypes.c_float()
        L.check(L.load().sred_relevance_mask_tool_2d(res.handle, L.fptr(a), w, h, block_width, block_height,
                                                     float(filter_constant), L.fptr(mask), ctypes.byref(nv),
                                                     ctypes.byref(thr)))
    finally:
        if own:
            res.close()
    return mask, nv.value, thr.value


def optimise_relevance_mask(image, block_width=3, block_height=3, step=1.0, max_iter=5000, max_steps=100_000_000,
                            resources=None):
    """OptimiseRelevanceMask_.run (OptimiseRelevanceMask_.java:30-287): the
    filter-constant range (last all-ones constant, first all-zeros constant)."""
    res, own = _ctx(resources)
    try:
        a, _ = prepare_input(image, max_iter, res)
        h, w = a.shape
        ones, zeros = ctypes.c_float(), ctypes.c_float()
        L.check(L.load().sred_optimise_relevance_mask_2d(res.handle, L.fptr(a), w, h, block_width, block_height,
                                                         float(step), int(max_steps), ctypes.byref(ones),
                                                         ctypes.byref(zeros)))
    finally:
        if own:
            res.close()
    return ones.value, zeros.value


def image_statistics_2d(image, block_width=3, block_height=3, max_iter=5000, resources=None):
    """What CalculateLocalMeans_ ("SReD: Image statistics") sets out to do:
    the local means and standard deviations of the stabilised image as a
    (2, H, W) stack.  The reference's own kernel setup cannot run (it fills
    $PATCH_SIZE$ into a kernel that reads $BLOCK_SIZE$ and asks for a kernel
    name the file does not define); the statistics are those of
    kernelGetLocalStatistics2D.cl, which that plugin loads."""
    if block_width % 2 == 0 or block_height % 2 == 0:
        raise ValueError("Patch dimensions must be odd (e.g., 3x3 or 5x5). Please try again.")
    res, own = _ctx(resources)
    try:
        a, _ = prepare_input(image, max_iter, res)
        h, w = a.shape
        st = clutils.get_local_statistics_2d(res, utils.InputImage2D(a, w, h, w * h), block_width, block_height)
        means, stds = st.local_means, st.local_stds
    finally:
        if own:
            res.close()
    return np.stack([np.asarray(means).reshape(h, w), np.asarray(stds).reshape(h, w)])


def relevance_mask_stack_2d(image, block_width=3, block_height=3, lower=0.0, upper=1.0, step=0.1, max_iter=5000,
                            resources=None):
    """RelevanceMaskStack2D_.run (RelevanceMaskStack2D_.java:31-301): the
    stabilised image and the relevance masks of the filter constants lower,
    lower + step, ... (float sums) up to the count the plugin computes.
    Returns (masks (n, H, W), constants (n,), prepared image)."""
    res, own = _ctx(resources)
    try:
        a, _ = prepare_input(image, max_iter, res)
        h, w = a.shape
        n = ctypes.c_int()
        rc = L.load().sred_relevance_mask_stack_2d(res.handle, L.fptr(a), w, h, block_width, block_height,
                                                   float(lower), float(upper), float(step), 0, None, None,
                                                   ctypes.byref(n))
        if n.value < 1:
            L.check(rc)  # raises: invalid block or constants
        masks = np.zeros((n.value, h, w), np.float32)
        consts = np.zeros(n.value, np.float32)
        L.check(L.load().sred_relevance_mask_stack_2d(res.handle, L.fptr(a), w, h, block_width, block_height,
                                                      float(lower), float(upper), float(step), n.value,
                                                      masks.ctypes.data, consts.ctypes.data, ctypes.byref(n)))
    finally:
        if own:
            res.close()
    return masks, consts, a
