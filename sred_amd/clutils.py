"""Mirror of SReD's ``CLUtils`` 2D entry points over the HIP library.

The Java CLUtils builds OpenCL programs and launches kernels through JOCL;
these functions keep its method names (snake_case), argument meaning, return
values and error behaviour, and call libsred_hip.so (include/sred.h) instead.
Unknown metric strings return None, as CLUtils.java:2193-2195 does.
"""
from dataclasses import dataclass

import numpy as np

from . import _lib as L
from .utils import InputImage2D, ReferenceBlock2D, RelevanceMask

GLOBAL_METRICS = ["Pearson's R", "SSIM", "NRMSE", "Abs. diff. of StdDevs"]          # GlobalRepetition2D_.java:20-25
BLOCK_METRICS = ["Pearson's R", "SSIM", "NRMSE (inverted)", "Abs. Diff. of StdDevs."]  # BlockRepetition2D_.java:18-23
PRECISIONS = {"fp32": L.FP32, "bf16x3": L.BF16X3}


@dataclass
class LocalStatistics:
    """CLUtils.CLLocalStatistics (CLUtils.java:84-154), host arrays."""

    local_means: np.ndarray
    local_stds: np.ndarray


def get_opencl_resources(device=0):
    """CLUtils.getOpenCLResources counterpart: a reusable HIP device context."""
    return L.Context(device)


def get_local_statistics_2d(resources, input_image, block_width, block_height):
    """CLUtils.getLocalStatistics2D (CLUtils.java:376-467)."""
    img = L.f32(input_image.image_array).reshape(-1)
    means = np.zeros(input_image.size, np.float32)
    stds = np.zeros(input_image.size, np.float32)
    L.check(L.load().sred_local_stats_2d(resources.handle, L.fptr(img), input_image.width, input_image.height,
                                         block_width, block_height, L.fptr(means), L.fptr(stds)))
    return LocalStatistics(means, stds)


def get_local_statistics_3d(resources, volume, block_width, block_height, block_depth):
    """CLUtils.getLocalStatistics3D (CLUtils.java:1030-1114) with its declared
    argument meaning: (means, stds) of a (depth, H, W) volume."""
    vol = L.f32(volume)
    d, h, w = vol.shape
    means = np.zeros_like(vol)
    stds = np.zeros_like(vol)
    L.check(L.load().sred_local_stats_3d(resources.handle, L.fptr(vol), w, h, d, block_width, block_height,
                                         block_depth, L.fptr(means), L.fptr(stds)))
    return LocalStatistics(means, stds)


def _metric_id(metric, block):
    rc = L.load().sred_metric_from_name(metric.encode(), 1 if block else 0)
    return None if rc == L.SRED_ERR_UNKNOWN_METRIC else L.check(rc)


def get_global_accumulators_2d(resources, metric_id, input_image, local_statistics, block_width, block_height,
                               filter_param, threshold, precision="bf16x3"):
    """Raw outputs of kernelGetGlobal*2D (similarity sum S, weight sum Wsum)."""
    img = L.f32(input_image.image_array).reshape(-1)
    S = np.zeros(input_image.size, np.float32)
    W = np.zeros(input_image.size, np.float32)
    L.check(L.load().sred_global_accum_2d(
        resources.handle, metric_id, PRECISIONS[precision], L.fptr(img), L.fptr(L.f32(local_statistics.local_means)),
        L.fptr(L.f32(local_statistics.local_stds)), input_image.width, input_image.height, block_width, block_height,
        filter_param, threshold, L.fptr(S), L.fptr(W)))
    return S, W


def calculate_global_repetition_map_2d(metric, input_image, block_width, block_height, block_size, local_statistics,
                                       relevance_mask, n_pixels, normalize_output, resources, precision="bf16x3"):
    """CLUtils.calculateGlobalRepetitionMap2D (CLUtils.java:2174-2196).

    precision selects the contraction: "bf16x3" (split-BF16 MFMA) or "fp32".
    Returns a float32 array of width*height, or None for an unknown metric.
    """
    mid = _metric_id(metric, block=False)
    if mid is None:
        return None
    if block_size != L.load().sred_block_size_2d(block_width, block_height):
        raise ValueError("block_size does not match the elliptical block of the given width/height")
    out = np.zeros(input_image.size, np.float32)
    img = L.f32(input_image.image_array).reshape(-1)
    L.check(L.load().sred_global_repetition_2d(
        resources.handle, mid, PRECISIONS[precision], L.fptr(img), L.fptr(L.f32(local_statistics.local_means)),
        L.fptr(L.f32(local_statistics.local_stds)), input_image.width, input_image.height, block_width, block_height,
        L.fptr(L.f32(relevance_mask.relevance_mask)), relevance_mask.relevance_constant,
        relevance_mask.relevance_threshold, relevance_mask.noise_mean_variance, n_pixels, int(bool(normalize_output)),
        L.fptr(out)))
    return out


def calculate_block_repetition_map_2d(metric, input_image, reference_block, relevance_constant, normalize_output,
                                      resources):
    """CLUtils.calculateBlockRepetitionMap2D (CLUtils.java:995-1010).

    Returns a float32 array of width*height, or None for an unknown metric.
    """
    mid = _metric_id(metric, block=True)
    if mid is None:
        return None
    out = np.zeros(input_image.size, np.float32)
    img = L.f32(input_image.image_array).reshape(-1)
    L.check(L.load().sred_block_repetition_2d(
        resources.handle, mid, L.fptr(img), input_image.width, input_image.height,
        L.fptr(L.f32(reference_block.pixels)), reference_block.width, reference_block.height, reference_block.mean,
        reference_block.std, relevance_constant, int(bool(normalize_output)), L.fptr(out)))
    return out


def calculate_block_repetition_map_3d(metric, input_image, reference_block, relevance_constant, normalize_output,
                                      resources):
    """CLUtils.calculateBlockRepetitionMap3D (CLUtils.java:1639-1653) with the
    block radii as intended (see sred_block_repetition_3d).

    Returns a float32 array of width*height*depth, or None for an unknown metric.
    """
    mid = _metric_id(metric, block=True)
    if mid is None:
        return None
    out = np.zeros(input_image.size, np.float32)
    img = L.f32(input_image.image_array).reshape(-1)
    rb = reference_block
    L.check(L.load().sred_block_repetition_3d(
        resources.handle, mid, L.fptr(img), input_image.width, input_image.height, input_image.depth,
        L.fptr(L.f32(rb.pixels)), rb.width, rb.height, rb.depth, rb.mean, rb.std, relevance_constant,
        int(bool(normalize_output)), L.fptr(out)))
    return out


def calculate_block_hu_map_2d(input_image, template, resources):
    """Hu template map (kernelGetPatchHu2D.cl:11-282; the reference has no Java
    caller for it): sqrt(hu1(template) * hu1(window)) per interior pixel.

    template is the (bh, bw) template rectangle as the kernel's patch_pixels
    (for example from utils.get_hu_template_2d).  Returns float32 w*h.
    """
    t = L.f32(template)
    bh, bw = t.shape
    out = np.zeros(input_image.size, np.float32)
    img = L.f32(input_image.image_array).reshape(-1)
    L.check(L.load().sred_block_hu_2d(resources.handle, L.fptr(img), input_image.width, input_image.height,
                                      L.fptr(t.reshape(-1)), bw, bh, L.fptr(out)))
    return out


def _rows_arg(unique_coords):
    if unique_coords is None:
        return None, 0, None
    r = np.ascontiguousarray(unique_coords, dtype=np.int32)
    return r.ctypes.data, r.size, r


def get_hu_invariants_2d(input_image, local_statistics, block_width, block_height, resources):
    """Per-pixel Hu invariant nu20 + nu02 of kernelGetHuMap.cl:146-172 (border 0)."""
    img = L.f32(input_image.image_array).reshape(-1)
    out = np.zeros(input_image.size, np.float32)
    L.check(L.load().sred_hu_invariants_2d(resources.handle, L.fptr(img), L.fptr(L.f32(local_statistics.local_means)),
                                           input_image.width, input_image.height, block_width, block_height,
                                           L.fptr(out)))
    return out


def calculate_global_hu_map_2d(input_image, local_statistics, block_width, block_height, resources,
                               unique_coords=None):
    """Global Hu map (kernelGetHuMap.cl:15-172; no Java caller in the reference):
    sum over interior j of |hu_j - hu_i| * (100^sim - 1) / 99.

    unique_coords: optional pixel indices y*w+x to compute (the kernel's
    uniqueStdCoords list); other pixels are 0.  Returns float32 w*h.
    """
    img = L.f32(input_image.image_array).reshape(-1)
    out = np.zeros(input_image.size, np.float32)
    rp, nr, _keep = _rows_arg(unique_coords)
    L.check(L.load().sred_global_hu_2d(resources.handle, L.fptr(img), L.fptr(L.f32(local_statistics.local_means)),
                                       L.fptr(L.f32(local_statistics.local_stds)), input_image.width,
                                       input_image.height, block_width, block_height, rp, nr, L.fptr(out)))
    return out


def calculate_phase_correlation_map_2d(input_image, local_statistics, block_width, block_height, resources,
                                       unique_coords=None):
    """Phase correlation map (kernelGetPhaseCorrelationMap.cl:13-179, which does
    not compile in the reference; textbook phase correlation on its blocks):
    sum over interior j of |peak displacement_ij| * w_ij.  Odd block sides
    3..11.  unique_coords as calculate_global_hu_map_2d.  Returns float32 w*h.
    """
    img = L.f32(input_image.image_array).reshape(-1)
    out = np.zeros(input_image.size, np.float32)
    rp, nr, _keep = _rows_arg(unique_coords)
    L.check(L.load().sred_phase_correlation_2d(
        resources.handle, L.fptr(img), L.fptr(L.f32(local_statistics.local_means)),
        L.fptr(L.f32(local_statistics.local_stds)), input_image.width, input_image.height, block_width, block_height,
        rp, nr, L.fptr(out)))
    return out
