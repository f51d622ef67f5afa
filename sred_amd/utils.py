"""Host-side mirror of SReD's ``Utils`` for the 2D path (Utils.java).

Same names (snake_case), argument meaning and error behaviour as the Java
methods; the arithmetic runs in libsred_hip.so's host C++ functions with Java
float semantics.  Variance stabilisation by the generalised Anscombe
transform runs on the GPU with the library's own simplex fit ("Simplex"
method); the "Quad/Octree" method raises NotImplementedError.
"""
import ctypes
from dataclasses import dataclass

import numpy as np

from . import _lib as L

EPSILON = np.float32(1e-7)  # Utils.EPSILON, Utils.java:20


@dataclass
class InputImage2D:
    """Utils.InputImage2D: the (normalised) float32 image, row-major."""

    image_array: np.ndarray
    width: int
    height: int
    size: int


@dataclass
class ReferenceBlock2D:
    """Utils.ReferenceBlock2D (Utils.java:108-216)."""

    pixels: np.ndarray
    width: int
    height: int
    radius_width: int
    radius_height: int
    size: int
    mean: float
    std: float


@dataclass
class RelevanceMask:
    """Utils.RelevanceMask (Utils.java:35-96)."""

    relevance_mask: np.ndarray
    relevance_constant: float
    relevance_threshold: float
    noise_mean_variance: float


def block_size_2d(block_width, block_height):
    """Number of pixels in the elliptical block (GlobalRepetition2D_.java:102-111)."""
    return L.check(L.load().sred_block_size_2d(block_width, block_height))


def get_gat_2d(pixels, gain, sigma, offset, resources=None):
    """VarianceStabilisingTransform2D_.getGAT (VarianceStabilisingTransform2D_.java:136-151).

    Returns a new float32 array (the Java method transforms in place).
    """
    arr = L.f32(np.array(pixels, dtype=np.float32, copy=True))
    own = resources is None
    res = L.Context(0) if own else resources
    try:
        L.check(L.load().sred_gat_2d(res.handle, L.fptr(arr.reshape(-1)), arr.size, float(gain), float(sigma),
                                     float(offset)))
    finally:
        if own:
            res.close()
    return arr


class GATMinimizer2D:
    """GATMinimizer2D (GATMinimizer2D.java:15-156): fits (gain, sigma, offset)
    so that the GAT-transformed image has unit variance in its (w/6) x (h/6)
    blocks.  run() uses the library's Nelder-Mead simplex in place of
    ImageJ's Minimizer (which is not restated)."""

    def __init__(self, pixels, width, height, gain, sigma, offset, max_iter, resources=None):
        self.pixels = L.f32(pixels).reshape(-1)
        self.width, self.height = int(width), int(height)
        self.gain, self.sigma, self.offset = float(gain), float(sigma), float(offset)
        self.max_iter = int(max_iter)
        self.error = None
        self._res = resources

    def _with_ctx(self, fn):
        own = self._res is None
        res = L.Context(0) if own else self._res
        try:
            return fn(res)
        finally:
            if own:
                res.close()

    def run(self):
        params = (ctypes.c_double * 3)()
        err = ctypes.c_double()

        def go(res):
            L.check(L.load().sred_gat_minimize_2d(res.handle, L.fptr(self.pixels), self.width, self.height, self.gain,
                                                  self.sigma, self.offset, self.max_iter, params,
                                                  ctypes.byref(err)))
        self._with_ctx(go)
        self.gain, self.sigma, self.offset = params[0], params[1], params[2]
        self.error = err.value

    def user_function(self, params, _v=0.0):
        """The objective at params = (gain, sigma, offset) (GATMinimizer2D.java:64-110)."""
        err = ctypes.c_double()

        def go(res):
            L.check(L.load().sred_gat_objective_2d(res.handle, L.fptr(self.pixels), self.width, self.height,
                                                   float(params[0]), float(params[1]), float(params[2]),
                                                   ctypes.byref(err)))
        self._with_ctx(go)
        return err.value


def get_input_image_2d(image, stabilise_noise_variance=False, normalize_output=True, gat_method="Simplex",
                       max_iter=5000, resources=None):
    """Utils.getInputImage2D (Utils.java:653-725, overload :746-775) for an array.

    Converts to float32; with stabilise_noise_variance, fits the GAT
    parameters (GATMinimizer2D with initial (1, 10, 100), Utils.java:695) and
    applies getGAT with them rounded to float (:697-701); then min-max
    normalises as (v-min)/(max-min+EPSILON).
    """
    arr = L.f32(np.array(image, dtype=np.float32, copy=True))
    if arr.ndim != 2:
        raise ValueError("expected a 2D image")
    h, w = arr.shape
    if stabilise_noise_variance:
        if gat_method != "Simplex":
            raise NotImplementedError("the Quad/Octree GAT method is not built")
        m = GATMinimizer2D(arr, w, h, 1, 10, 100, max_iter, resources)
        m.run()
        arr = get_gat_2d(arr, np.float32(m.gain), np.float32(m.sigma), np.float32(m.offset), resources).reshape(h, w)
    if normalize_output:
        L.check(L.load().sred_normalize_input_2d(L.fptr(arr), arr.size))
    return InputImage2D(arr, w, h, w * h)


def get_reference_block_2d(block):
    """Utils.getReferenceBlock2D (Utils.java:557-639) for a 2D array.

    Returns None for even block dimensions, as the reference does.
    """
    b = L.f32(block)
    bh, bw = b.shape
    if bw % 2 == 0 or bh % 2 == 0:
        return None
    k = block_size_2d(bw, bh)
    px = np.zeros(k, np.float32)
    m, s = ctypes.c_float(), ctypes.c_float()
    k = L.check(L.load().sred_reference_block_2d(L.fptr(b), bw, bh, L.fptr(px), ctypes.pointer(m), ctypes.pointer(s)))
    return ReferenceBlock2D(px, bw, bh, bw // 2, bh // 2, k, m.value, s.value)


def get_mean_noise_var_2d(ref_pixels, w, h):
    """Utils.getMeanNoiseVar2D (Utils.java:790-837)."""
    a = L.f32(ref_pixels).reshape(-1)
    out = ctypes.c_float()
    L.check(L.load().sred_noise_var_2d(L.fptr(a), w, h, ctypes.pointer(out)))
    return out.value


def get_relevance_mask(image_array, image_width, image_height, block_radius_width, block_radius_height, local_stds,
                       relevance_constant):
    """Utils.getRelevanceMask (Utils.java:944-968)."""
    nv = get_mean_noise_var_2d(image_array, image_width, image_height)
    stds = L.f32(local_stds).reshape(-1)
    mask = np.zeros(image_width * image_height, np.float32)
    thr, npx = ctypes.c_float(), ctypes.c_float()
    L.check(L.load().sred_relevance_mask_2d(L.fptr(stds), image_width, image_height, block_radius_width,
                                            block_radius_height, nv, relevance_constant, L.fptr(mask),
                                            ctypes.pointer(thr), ctypes.pointer(npx)))
    return RelevanceMask(mask, float(np.float32(relevance_constant)), thr.value, nv)


def count_relevant_pixels(relevance_mask, image_width, image_height, block_radius_width, block_radius_height):
    """nPixels of GlobalRepetition2D_.java:133-140 (float count of mask > 0 on the interior)."""
    m = np.asarray(relevance_mask, np.float32).reshape(image_height, image_width)
    inner = m[block_radius_height:image_height - block_radius_height, block_radius_width:image_width - block_radius_width]
    return float(min(int((inner > 0).sum()), 1 << 24))


def apply_mask_2d(image_array, image_width, image_height, mask):
    """Utils.applyMask2D (Utils.java:915-925): in place, returns the array."""
    a = image_array
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"]
    m = L.f32(mask).reshape(-1)
    L.check(L.load().sred_apply_mask_2d(L.fptr(a.reshape(-1)), image_width, image_height, L.fptr(m)))
    return a


def normalize_image_2d(image_array, image_width, image_height, border_width, border_height, mask=None):
    """Utils.normalizeImage2D (Utils.java:985-1031): returns a normalised copy.

    Raises NonBinaryMaskError (a ValueError) for a non-binary mask, as the
    reference throws IllegalArgumentException.
    """
    out = L.f32(np.array(image_array, dtype=np.float32, copy=True))
    m = None if mask is None else L.f32(mask).reshape(-1)
    L.check(L.load().sred_normalize_map_2d(L.fptr(out.reshape(-1)), image_width, image_height, border_width,
                                           border_height, L.fptr(m)))
    return out


def get_hu_template_2d(block):
    """Template rectangle for the Hu map: the window steps of
    kernelGetPatchHu2D.cl:127-182 (mean over the square minus corners / K,
    subtraction and min-max over the ellipse) applied to a (bh, bw) block.
    Even sizes raise ValueError (the reference's block commands reject them)."""
    b = L.f32(block)
    bh, bw = b.shape
    out = np.zeros_like(b)
    L.check(L.load().sred_hu_prepare_template_2d(L.fptr(b.reshape(-1)), bw, bh, L.fptr(out.reshape(-1))))
    return out
