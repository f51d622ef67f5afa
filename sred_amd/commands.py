"""The two 2D plugin commands on the HIP path.

``global_repetition_2d`` follows GlobalRepetition2D_.run
(GlobalRepetition2D_.java:35-157) and ``block_repetition_2d`` follows
BlockRepetition2D_.run (BlockRepetition2D_.java:32-157, single image) with
the dialog parameters as keyword arguments and the same defaults (block 5x5,
relevance constant 0, normalise on).  The ImageJ dialog, window lookup,
display and variance stabilisation stay outside this library.
"""
import numpy as np

from . import clutils, utils
from . import _lib as L


def global_repetition_2d(image, block_width=5, block_height=5, relevance_constant=0.0, metric="Pearson's R",
                         normalize_output=True, precision="bf16x3", resources=None):
    """Repetition map of every block against every other block (float32, HxW)."""
    if block_width % 2 == 0 or block_height % 2 == 0:  # GlobalRepetition2D_.java:86-89
        raise ValueError("Block dimensions must be odd (e.g., 3x3 or 5x5). Please try again.")
    own = resources is None
    res = clutils.get_opencl_resources() if own else resources
    try:
        brw, brh = block_width // 2, block_height // 2
        block_size = utils.block_size_2d(block_width, block_height)
        inp = utils.get_input_image_2d(image, False, True)
        stats = clutils.get_local_statistics_2d(res, inp, block_width, block_height)
        mask = utils.get_relevance_mask(inp.image_array, inp.width, inp.height, brw, brh, stats.local_stds,
                                        relevance_constant)
        n_pixels = utils.count_relevant_pixels(mask.relevance_mask, inp.width, inp.height, brw, brh)
        out = clutils.calculate_global_repetition_map_2d(metric, inp, block_width, block_height, block_size, stats,
                                                         mask, n_pixels, normalize_output, res, precision)
        return None if out is None else out.reshape(inp.height, inp.width)
    finally:
        if own:
            res.close()


def block_repetition_2d(block, image, relevance_constant=0.0, metric="Pearson's R", normalize_output=True,
                        resources=None):
    """Repetition map of one reference block against every window (float32, HxW)."""
    ref = utils.get_reference_block_2d(block)
    if ref is None:
        raise ValueError("Block dimensions must be odd (e.g., 3x3 or 5x5). Please try again.")
    own = resources is None
    res = clutils.get_opencl_resources() if own else resources
    try:
        inp = utils.get_input_image_2d(image, False, True)
        out = clutils.calculate_block_repetition_map_2d(metric, inp, ref, relevance_constant, normalize_output, res)
        return None if out is None else out.reshape(inp.height, inp.width)
    finally:
        if own:
            res.close()


def global_repetition_3d(volume, block_width=5, block_height=5, block_depth=5, relevance_constant=0.0,
                         metric="Pearson's R", normalize_output=True, precision="bf16x3", resources=None,
                         stabilise_noise_variance=False, gat_method="Simplex"):
    """GlobalRepetition3D_.run (GlobalRepetition3D_.java:35-160) on a (depth,
    H, W) volume with one of its metrics ("Pearson's R", "SSIM", "NRMSE",
    "Abs. diff. of StdDevs"): the input step of Utils.getInputImage3D
    (utils.get_input_image_3d: normalised as a whole), then the map built as
    intended where the reference does not run (see
    sred_global_repetition_3d); returns a float32 (depth, H, W) map, or None
    for another metric name (CLUtils.calculateGlobalRepetitionMap3D returns
    null)."""
    if block_width % 2 == 0 or block_height % 2 == 0 or block_depth % 2 == 0:  # :80-84
        raise ValueError("Block dimensions must be odd. Please try again.")
    if block_depth < 3:  # :86-90
        raise ValueError("Block depth must be at least 3. Please try again.")
    inp = utils.get_input_image_3d(volume, stabilise_noise_variance, gat_method, True, resources)
    if inp is None:
        raise ValueError("Image must have at least 3 slices. Please try again.")
    m = L.load().sred_metric_from_name(metric.encode(), 0)
    if m < 0:
        return None
    vol = inp.image_array
    d, h, w = vol.shape
    out = np.zeros_like(vol)
    own = resources is None
    res = clutils.get_opencl_resources() if own else resources
    try:
        L.check(L.load().sred_global_repetition_3d(res.handle, m, clutils.PRECISIONS[precision], L.fptr(vol),
                                                   w, h, d, block_width, block_height, block_depth,
                                                   float(relevance_constant), int(bool(normalize_output)),
                                                   L.fptr(out)))
    finally:
        if own:
            res.close()
    return out


def block_repetition_timelapse_2d(block, stack, relevance_constant=0.0, metric="Pearson's R",
                                  normalize_output=True, stabilise_noise_variance=False, gat_method="Simplex",
                                  resources=None):
    """The time-lapse branch of BlockRepetition2D_ (BlockRepetition2D_.java:95-137):
    one repetition map per frame of a (depth, H, W) stack (depth >= 3), as a
    float32 (depth, H, W) array.  The stack goes through getInputImage3D
    (:97) first: with stabilisation and "Quad/Octree" the octree GAT of the
    whole stack (utils.get_input_image_3d), with "Simplex" nothing (the
    reference discards that transform); the entry then normalises the stack
    and each frame."""
    ref = utils.get_reference_block_2d(block)
    if ref is None:
        raise ValueError("Block dimensions must be odd (e.g., 3x3 or 5x5). Please try again.")
    st = L.f32(stack)
    if st.ndim != 3:
        raise ValueError("expected a (depth, height, width) stack")
    d, h, w = st.shape
    m = L.load().sred_metric_from_name(metric.encode(), 1)
    if m < 0:
        return None
    out = np.zeros_like(st)
    own = resources is None
    res = clutils.get_opencl_resources() if own else resources
    try:
        if stabilise_noise_variance and gat_method == "Quad/Octree" and d >= 3:
            st = L.f32(utils.get_input_image_3d(st, True, gat_method, False, res).image_array.reshape(d, h, w))
        L.check(L.load().sred_block_timelapse_2d(res.handle, m, L.fptr(st), w, h, d, L.fptr(L.f32(ref.pixels)),
                                                 ref.width, ref.height, ref.mean, ref.std, float(relevance_constant),
                                                 int(bool(normalize_output)), L.fptr(out)))
    finally:
        if own:
            res.close()
    return out


def global_map_device(ctx, metric_id, precision_id, d_img, width, height, block_width, block_height,
                      relevance_constant, normalize_output, d_out, d_means=0, d_stds=0, d_mask=0, d_S=0, d_W=0):
    """sred_global_map_2d_dev on device pointers (ints, e.g. torch data_ptr())."""
    L.check(L.load().sred_global_map_2d_dev(ctx.handle, metric_id, precision_id, d_img, width, height, block_width,
                                            block_height, relevance_constant, int(bool(normalize_output)), d_out,
                                            d_means or None, d_stds or None, d_mask or None, d_S or None, d_W or None))


def block_map_device(ctx, metric_id, d_img, width, height, reference_block, relevance_constant, normalize_output,
                     d_out):
    """sred_block_map_2d_dev on device pointers; the template stays a host array."""
    L.check(L.load().sred_block_map_2d_dev(ctx.handle, metric_id, d_img, width, height,
                                           L.fptr(L.f32(reference_block.pixels)), reference_block.width,
                                           reference_block.height, reference_block.mean, reference_block.std,
                                           relevance_constant, int(bool(normalize_output)), d_out))


def global_batch_device(ctx, metric_id, problems, relevance_constant=0.0, normalize_output=True):
    """sred_global_batch_2d: problems is a list of (d_img, width, height, block_width, block_height, d_out)
    device-pointer tuples; every map is written into its d_out."""
    arr = (L.GlobalProblem * len(problems))()
    for i, (d_img, w, h, bw, bh, d_out) in enumerate(problems):
        arr[i] = L.GlobalProblem(d_img, w, h, bw, bh, d_out)
    L.check(L.load().sred_global_batch_2d(ctx.handle, metric_id, len(problems), arr, relevance_constant,
                                          int(bool(normalize_output))))


def block_repetition_3d(block, volume, relevance_constant=0.0, metric="Pearson's R", normalize_output=True,
                        resources=None, stabilise_noise_variance=False, gat_method="Simplex"):
    """BlockRepetition3D_.run (BlockRepetition3D_.java:32-115) on a (depth, H,
    W) volume and a (depth, H, W) reference block: Utils.getReferenceBlock3D,
    the input step of Utils.getInputImage3D, then
    CLUtils.calculateBlockRepetitionMap3D with the block radii as intended.
    Returns a float32 (depth, H, W) map, or None for an unknown metric."""
    ref = utils.get_reference_block_3d(block)
    inp = utils.get_input_image_3d(volume, stabilise_noise_variance, gat_method, True, resources)
    if inp is None:
        raise ValueError("Image must have at least 3 slices. Please try again.")
    if ref.width > inp.width or ref.height > inp.height or ref.depth > inp.depth:  # :97-100
        raise ValueError("Block dimensions must be smaller than image dimensions.")
    own = resources is None
    res = clutils.get_opencl_resources() if own else resources
    try:
        out = clutils.calculate_block_repetition_map_3d(metric, inp, ref, relevance_constant, normalize_output, res)
        return None if out is None else out.reshape(inp.depth, inp.height, inp.width)
    finally:
        if own:
            res.close()


def variance_stabilising_transform_3d(volume, gain=1.0, offset=10.0, sigma=100.0, max_iter=5000, resources=None):
    """The VarianceStabilisingTransform3D_ plugin (VarianceStabilisingTransform3D_
    .java:28-110) on a (depth, H, W) volume: GATMinimizer3D from the dialog's
    initial (gain, offset, noise std) and the GAT of getGAT with the fitted
    (double) parameters on every voxel.  Returns (float32 volume, (gain,
    sigma, offset))."""
    vol = L.f32(volume)
    if vol.ndim != 3 or vol.shape[0] < 2:  # :47-50
        raise ValueError("Image is not 3D. Please open a 3D image and try again.")
    d, h, w = vol.shape
    own = resources is None
    res = clutils.get_opencl_resources() if own else resources
    try:
        m = utils.GATMinimizer3D(vol, w, h, d, gain, sigma, offset, max_iter, res)
        m.run()
        out = utils.get_gat_2d(vol, m.gain, m.sigma, m.offset, res).reshape(d, h, w)
        return out, (m.gain, m.sigma, m.offset)
    finally:
        if own:
            res.close()


def _clip_roi(roi, width, height):
    """The rectangle ip.getRoi() returns for roi = (x, y, w, h): ImageJ's
    ImageProcessor.setRoi keeps its intersection with the image, and an
    empty intersection becomes the empty rectangle (0, 0, 0, 0)."""
    x, y, w, h = (int(v) for v in roi)
    x0, y0 = max(x, 0), max(y, 0)
    x1, y1 = min(x + w, width), min(y + h, height)
    if x1 <= x0 or y1 <= y0:
        return 0, 0, 0, 0
    return x0, y0, x1 - x0, y1 - y0


def _roi_mean_std(values):
    """VarianceStabilisingTransform2D_.meanAndStdDev (:116-133): double sums in
    order, mean and sqrt(sq/n - mean^2); (0, 0) for no values."""
    a = np.asarray(values, np.float64).reshape(-1)
    if a.size == 0:
        return 0.0, 0.0
    s = sq = 0.0
    for v in a.tolist():
        s += v
        sq += v * v
    mean = s / a.size
    return mean, float(np.sqrt(sq / a.size - mean * mean))


def variance_stabilising_transform_2d(image, gain=1.0, offset=10.0, sigma=100.0, max_iter=5000, roi=None,
                                      resources=None):
    """The VarianceStabilisingTransform2D_ plugin (VarianceStabilisingTransform2D_
    .java:24-110) on a 2D image: with roi = (x, y, w, h) the initial offset and
    noise std come from the ROI's pixels as ImageProcessor.getPixel returns
    them (the value for 8- and 16-bit images; for a 32-bit image ImageJ's
    FloatProcessor returns the float's raw bits as an int, and so does this
    mirror for float input); then GATMinimizer2D from (gain, sigma, offset)
    and getGAT with the fitted doubles.  Returns (float32 image, (gain,
    sigma, offset))."""
    arr = np.asarray(image)
    if arr.ndim != 2:
        raise ValueError("expected a 2D image")
    if roi is not None:
        x, y, w, h = _clip_roi(roi, arr.shape[1], arr.shape[0])
        vals = arr[y:y + h, x:x + w]
        if np.issubdtype(vals.dtype, np.floating):
            vals = np.ascontiguousarray(vals, np.float32).view(np.int32)
        offset, sigma = _roi_mean_std(vals)
    px = L.f32(arr)
    hh, ww = px.shape
    own = resources is None
    res = clutils.get_opencl_resources() if own else resources
    try:
        m = utils.GATMinimizer2D(px, ww, hh, gain, sigma, offset, max_iter, res)
        m.run()
        out = utils.get_gat_2d(px, m.gain, m.sigma, m.offset, res).reshape(hh, ww)
        return out, (m.gain, m.sigma, m.offset)
    finally:
        if own:
            res.close()


def quadtree_stabilise_2d(image, min_size=4, alpha=0.01, max_iterations=50, var_alpha=0.75, resources=None):
    """The QuadTree_ plugin (QuadTree_.java:28-100) on a 2D image: the tree,
    robust means (max_iterations), LTS variances (var_alpha), the regression
    and applyGATtree with its double (g0, eDC).  Returns (float32 image,
    (g0, eDC), leaf rectangles (x, y, w, h) -- its overlay)."""
    arr = L.f32(np.array(image, dtype=np.float32, copy=True))
    if arr.ndim != 2:
        raise ValueError("expected a 2D image")
    h, w = arr.shape
    own = resources is None
    res = clutils.get_opencl_resources() if own else resources
    try:
        qt = utils.QuadTree(w, h, min_size, alpha, res)
        qt.build_tree(arr)
        qt.calculate_robust_means(arr, max_iterations)
        qt.calculate_lts_variances(arr, var_alpha)
        g0, edc = qt.perform_linear_regression(qt.collect_mean_variance_pairs())
        out = qt.apply_gat_tree(arr, w * h, g0, edc).reshape(h, w)
        return out, (g0, edc), qt.leaves()
    finally:
        if own:
            res.close()


def octree_stabilise_3d(volume, min_size=4, alpha=0.01, max_iterations=50, var_alpha=0.75, resources=None):
    """The OcTree_ plugin (OcTree_.java:30-126) on a (depth, H, W) volume of at
    least 3 slices: the tree, robust means (max_iterations), LTS variances
    with the fraction 0.75 (the plugin reads var_alpha but passes 0.75f,
    :88), the float (g0, eDC) and applyGATtree over every voxel.  Returns
    (float32 volume, (g0, eDC), leaf boxes (x, y, z, w, h, d))."""
    arr = L.f32(np.array(volume, dtype=np.float32, copy=True))
    if arr.ndim != 3 or arr.shape[0] < 3:  # :64-67
        raise ValueError("Image must have at least 3 slices. Please try again.")
    d, h, w = arr.shape
    own = resources is None
    res = clutils.get_opencl_resources() if own else resources
    try:
        oc = utils.OcTree(w, h, d, min_size, alpha, res)
        oc.build_tree(arr)
        oc.calculate_robust_means(arr, max_iterations)
        oc.calculate_lts_variances(arr, 0.75)
        g0, edc = oc.perform_linear_regression(oc.collect_mean_variance_pairs())
        out = oc.apply_gat_tree(arr, w * h * d, g0, edc).reshape(d, h, w)
        return out, (g0, edc), oc.leaves()
    finally:
        if own:
            res.close()
