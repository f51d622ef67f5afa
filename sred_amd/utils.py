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
class InputImage3D:
    """Utils.InputImage3D: the (normalised) float32 volume, slice z at z*w*h."""

    image_array: np.ndarray
    width: int
    height: int
    depth: int
    size: int


@dataclass
class ReferenceBlock3D:
    """Utils.ReferenceBlock3D: the ellipsoid's voxels of the block and their statistics."""

    pixels: np.ndarray
    width: int
    height: int
    depth: int
    radius_width: int
    radius_height: int
    radius_depth: int
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


class GATMinimizer3D(GATMinimizer2D):
    """GATMinimizer3D (GATMinimizer3D.java:19-158): the same fit over the
    volume's blocks (sides max(n/6, 1), halved along axes with fewer than 2
    blocks), squares in double, NaN for gain <= 0, sigma < 0 or offset < 0.
    run() uses the library's Nelder-Mead simplex in place of ImageJ's
    Minimizer."""

    def __init__(self, pixels, width, height, depth, gain, sigma, offset, max_iter, resources=None):
        super().__init__(pixels, width, height, gain, sigma, offset, max_iter, resources)
        self.depth = int(depth)

    def run(self):
        params = (ctypes.c_double * 3)()
        err = ctypes.c_double()

        def go(res):
            L.check(L.load().sred_gat_minimize_3d(res.handle, L.fptr(self.pixels), self.width, self.height,
                                                  self.depth, self.gain, self.sigma, self.offset, self.max_iter,
                                                  params, ctypes.byref(err)))
        self._with_ctx(go)
        self.gain, self.sigma, self.offset = params[0], params[1], params[2]
        self.error = err.value

    def user_function(self, params, _v=0.0):
        """The objective at params = (gain, sigma, offset) (GATMinimizer3D.java:79-157)."""
        err = ctypes.c_double()

        def go(res):
            L.check(L.load().sred_gat_objective_3d(res.handle, L.fptr(self.pixels), self.width, self.height,
                                                   self.depth, float(params[0]), float(params[1]),
                                                   float(params[2]), ctypes.byref(err)))
        self._with_ctx(go)
        return err.value


class QuadTree:
    """QuadTree_ (QuadTree_.java:119-427): the Quad/Octree estimate of the GAT
    parameters.  The device computes the tree, the robust means and the LTS
    variances in one call (sred_quadtree_gat_2d) when the results are first
    needed, with the parameters given to the constructor and to
    calculate_robust_means / calculate_lts_variances."""

    def __init__(self, image_width, image_height, min_size, alpha, resources=None):
        self.image_width, self.image_height = int(image_width), int(image_height)
        self.min_size, self.alpha = int(min_size), float(np.float32(alpha))
        self.max_iterations, self.var_alpha = 50, float(np.float32(0.75))
        self._pixels = None
        self._res = resources
        self._out = None

    def build_tree(self, image_data):
        self._pixels = L.f32(image_data).reshape(-1)
        if self._pixels.size != self.image_width * self.image_height:
            raise ValueError("image size does not match the tree")
        self._out = None

    def calculate_robust_means(self, image_data, max_iterations):
        self.max_iterations = int(max_iterations)
        self._out = None

    def calculate_lts_variances(self, image_data, alpha):
        self.var_alpha = float(np.float32(alpha))
        self._out = None

    def _run(self):
        if self._out is not None:
            return self._out
        if self._pixels is None:
            raise RuntimeError("build_tree first")
        w, h = self.image_width, self.image_height
        cap = max(16, (w * h) // 2)
        rects = np.zeros((cap, 4), np.int32)
        means = np.zeros(cap, np.float32)
        var = np.zeros(cap, np.float32)
        ge = (ctypes.c_double * 2)()
        n = ctypes.c_int(0)
        own = self._res is None
        res = L.Context(0) if own else self._res
        try:
            rc = L.load().sred_quadtree_gat_2d(res.handle, L.fptr(self._pixels), w, h, self.min_size, self.alpha,
                                               self.max_iterations, self.var_alpha, ge, ctypes.byref(n), cap,
                                               rects.ctypes.data, L.fptr(means), L.fptr(var))
        finally:
            if own:
                res.close()
        k = n.value
        if rc < 0 and not (rc == L.SRED_ERR_ARG and 0 < k < 3):
            L.check(rc)  # raises
        self._out = (rects[:k].copy(), means[:k].copy(), var[:k].copy())
        return self._out

    def leaves(self):
        """Leaf rectangles (x, y, w, h) in pre-order (createOverlay's rectangles)."""
        return self._run()[0]

    def collect_mean_variance_pairs(self):
        """[[robust mean, LTS variance], ...] over the leaves in pre-order (as doubles)."""
        _, m, v = self._run()
        return [[float(a), float(b)] for a, b in zip(m, v)]

    def perform_linear_regression(self, pairs):
        """(g0, eDC) = slope and intercept of SimpleRegression over the pairs;
        ValueError below 3 pairs (SimpleRegression.regress throws)."""
        p = np.asarray(pairs, np.float64).reshape(-1, 2)
        if len(p) < 3:
            raise ValueError("not enough data (at least 3 points) for a regression")
        x, y = np.ascontiguousarray(p[:, 0]), np.ascontiguousarray(p[:, 1])
        s, i = ctypes.c_double(), ctypes.c_double()
        dp = ctypes.POINTER(ctypes.c_double)
        L.check(L.load().sred_simple_regression(x.ctypes.data_as(dp), y.ctypes.data_as(dp), len(p), ctypes.byref(s),
                                                ctypes.byref(i)))
        return [s.value, i.value]

    def apply_gat_tree(self, image_data, n_pixels, g0, edc):
        """QuadTree_.applyGATtree: a new float32 array of the first n_pixels."""
        out = L.f32(image_data).reshape(-1)[:n_pixels].copy()
        own = self._res is None
        res = L.Context(0) if own else self._res
        try:
            L.check(L.load().sred_gat_tree_2d(res.handle, L.fptr(out), out.size, float(g0), float(edc)))
        finally:
            if own:
                res.close()
        return out


def f_quantile(d1, d2, p):
    """Inverse CDF of the F(d1, d2) distribution (FDistribution.inverseCumulativeProbability)."""
    out = ctypes.c_double()
    L.check(L.load().sred_f_quantile(float(d1), float(d2), float(p), ctypes.byref(out)))
    return out.value


def get_input_image_2d(image, stabilise_noise_variance=False, normalize_output=True, gat_method="Simplex",
                       max_iter=5000, resources=None):
    """Utils.getInputImage2D (Utils.java:653-725, overload :746-775) for an array.

    Converts to float32; with stabilise_noise_variance and gat_method
    "Simplex", fits the GAT parameters (GATMinimizer2D with initial
    (1, 10, 100), Utils.java:695) and applies getGAT with them rounded to
    float (:697-701); with "Quad/Octree", estimates (g0, eDC) with
    QuadTree_(w, h, 4, 0.01f), 50 robust-mean iterations and LTS fraction
    0.75f and applies applyGATtree (:686-703); another method string
    stabilises nothing (the reference compares the strings by identity).
    Then min-max normalises as (v-min)/(max-min+EPSILON).
    """
    arr = L.f32(np.array(image, dtype=np.float32, copy=True))
    if arr.ndim != 2:
        raise ValueError("expected a 2D image")
    h, w = arr.shape
    if stabilise_noise_variance and gat_method == "Simplex":
        m = GATMinimizer2D(arr, w, h, 1, 10, 100, max_iter, resources)
        m.run()
        arr = get_gat_2d(arr, np.float32(m.gain), np.float32(m.sigma), np.float32(m.offset), resources).reshape(h, w)
    if stabilise_noise_variance and gat_method == "Quad/Octree":
        qt = QuadTree(w, h, 4, 0.01, resources)
        qt.build_tree(arr)
        qt.calculate_robust_means(arr, 50)
        qt.calculate_lts_variances(arr, 0.75)
        g0, edc = qt.perform_linear_regression(qt.collect_mean_variance_pairs())
        arr = qt.apply_gat_tree(arr, w * h, g0, edc).reshape(h, w)
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


class OcTree:
    """OcTree_ (OcTree_.java:136-486): the octree estimate of the GAT
    parameters of a volume.  The device computes the tree, the robust means
    and the LTS variances in one call (sred_octree_gat_3d) when the results
    are first needed; perform_linear_regression returns float values, as
    the reference's does."""

    def __init__(self, image_width, image_height, image_depth, min_size, alpha, resources=None):
        self.image_width, self.image_height, self.image_depth = int(image_width), int(image_height), int(image_depth)
        self.min_size, self.alpha = int(min_size), float(np.float32(alpha))
        self.max_iterations, self.var_alpha = 50, float(np.float32(0.75))
        self._voxels = None
        self._res = resources
        self._out = None

    def build_tree(self, image_data):
        self._voxels = L.f32(image_data).reshape(-1)
        if self._voxels.size != self.image_width * self.image_height * self.image_depth:
            raise ValueError("volume size does not match the tree")
        self._out = None

    def calculate_robust_means(self, image_data, max_iterations):
        self.max_iterations = int(max_iterations)
        self._out = None

    def calculate_lts_variances(self, image_data, alpha):
        self.var_alpha = float(np.float32(alpha))
        self._out = None

    def _run(self):
        if self._out is not None:
            return self._out
        if self._voxels is None:
            raise RuntimeError("build_tree first")
        w, h, d = self.image_width, self.image_height, self.image_depth
        cap = max(16, (w * h * d) // 2)
        boxes = np.zeros((cap, 6), np.int32)
        means = np.zeros(cap, np.float32)
        var = np.zeros(cap, np.float32)
        ge = (ctypes.c_double * 2)()
        n = ctypes.c_int(0)
        own = self._res is None
        res = L.Context(0) if own else self._res
        try:
            rc = L.load().sred_octree_gat_3d(res.handle, L.fptr(self._voxels), w, h, d, self.min_size, self.alpha,
                                             self.max_iterations, self.var_alpha, ge, ctypes.byref(n), cap,
                                             boxes.ctypes.data, means.ctypes.data, var.ctypes.data)
        finally:
            if own:
                res.close()
        k = n.value
        if rc < 0 and not (rc == L.SRED_ERR_ARG and 0 < k < 3):
            L.check(rc)  # raises
        self._out = (boxes[:k].copy(), means[:k].copy(), var[:k].copy())
        return self._out

    def leaves(self):
        """Leaf boxes (x, y, z, w, h, d) in pre-order."""
        return self._run()[0]

    def collect_mean_variance_pairs(self):
        """[[robust mean, LTS variance], ...] over the leaves in pre-order (as doubles)."""
        _, m, v = self._run()
        return [[float(a), float(b)] for a, b in zip(m, v)]

    def perform_linear_regression(self, pairs):
        """[g0, eDC] as float values (OcTree_.java:458-469); ValueError below 3 pairs."""
        g0, edc = QuadTree.perform_linear_regression(self, pairs)
        return [float(np.float32(g0)), float(np.float32(edc))]

    def apply_gat_tree(self, image_data, n_pixels, g0, edc):
        """OcTree_.applyGATtree: a new float32 array of the first n_pixels voxels."""
        out = L.f32(image_data).reshape(-1)[:n_pixels].copy()
        own = self._res is None
        res = L.Context(0) if own else self._res
        try:
            L.check(L.load().sred_gat_tree_3d(res.handle, L.fptr(out), out.size, float(g0), float(edc)))
        finally:
            if own:
                res.close()
        return out


def get_input_image_3d(volume, stabilise_noise_variance=False, gat_method="Simplex", normalize_output=True,
                       resources=None):
    """Utils.getInputImage3D (Utils.java:1167-1263) for a (depth, H, W) array.

    Returns None below 3 slices (:1192-1196).  The "Simplex" stabilisation
    leaves the volume unchanged: the reference writes the transformed
    slices into a stack array that it then discards (:1227-1229), so only
    its (unused) fitted parameters differ.  "Quad/Octree" estimates (g0, eDC)
    with OcTree_(w, h, d, 4, 0.01f), 50 robust-mean iterations and LTS
    fraction 0.75f and applies applyGATtree to every voxel (the reference
    passes width*height, :1253, and the one-slice array it gets back
    overruns the volume buffers downstream).  Then min-max normalises the
    whole volume as (v-min)/(max-min+EPSILON) (Utils.normalizeArray,
    :1257-1260).
    """
    arr = L.f32(np.array(volume, dtype=np.float32, copy=True))
    if arr.ndim != 3:
        raise ValueError("expected a (depth, height, width) volume")
    d, h, w = arr.shape
    if d < 3:
        return None
    if stabilise_noise_variance and gat_method == "Quad/Octree":
        oc = OcTree(w, h, d, 4, 0.01, resources)
        oc.build_tree(arr)
        oc.calculate_robust_means(arr, 50)
        oc.calculate_lts_variances(arr, 0.75)
        g0, edc = oc.perform_linear_regression(oc.collect_mean_variance_pairs())
        arr = oc.apply_gat_tree(arr, w * h * d, g0, edc).reshape(d, h, w)
    if normalize_output:
        L.check(L.load().sred_normalize_input_2d(L.fptr(arr), arr.size))
    return InputImage3D(arr, w, h, d, w * h * d)


def get_reference_block_3d(block):
    """Utils.getReferenceBlock3D (Utils.java:1047-1155) for a (depth, H, W) block.

    Raises ValueError for even sides or fewer than 3 slices, as the
    reference throws IllegalArgumentException (:1066-1076).
    """
    b = L.f32(block)
    if b.ndim != 3:
        raise ValueError("expected a (depth, height, width) block")
    bd, bh, bw = b.shape
    if bw % 2 == 0 or bh % 2 == 0 or bd % 2 == 0:
        raise ValueError("Block dimensions must be odd.")
    if bd < 3:
        raise ValueError("Reference block must have at least 3 slices.")
    k = L.check(L.load().sred_block_size_3d(bw, bh, bd))
    px = np.zeros(k, np.float32)
    m, s = ctypes.c_float(), ctypes.c_float()
    k = L.check(L.load().sred_reference_block_3d(L.fptr(b), bw, bh, bd, L.fptr(px), ctypes.pointer(m),
                                                 ctypes.pointer(s)))
    return ReferenceBlock3D(px, bw, bh, bd, bw // 2, bh // 2, bd // 2, k, m.value, s.value)


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
