"""sred_amd -- SReD's 2D block-correlation path on MI355X (gfx950).

Hand-written HIP kernels in libsred_hip.so behind the C ABI of
include/sred.h; this package is the host-side mirror of SReD's CLUtils and
Utils interfaces for that path (ctypes), plus the two 2D commands.
"""
from . import _lib
from ._lib import (ABSDIFF, BF16X3, FP32, NRMSE, PEARSON, SSIM, Context, NonBinaryMaskError, SredError,
                   load)
from .clutils import (BLOCK_METRICS, GLOBAL_METRICS, calculate_block_hu_map_2d, calculate_block_repetition_map_2d,
                      calculate_block_repetition_map_3d,
                      calculate_global_hu_map_2d, calculate_global_repetition_map_2d,
                      calculate_phase_correlation_map_2d, get_hu_invariants_2d, get_local_statistics_2d,
                      get_opencl_resources)
from .commands import (block_repetition_2d, block_repetition_3d, block_repetition_timelapse_2d, global_repetition_2d,
                       global_repetition_3d, octree_stabilise_3d, quadtree_stabilise_2d,
                       variance_stabilising_transform_2d, variance_stabilising_transform_3d)
from .utils import (GATMinimizer2D, GATMinimizer3D, InputImage2D, InputImage3D, OcTree, QuadTree, ReferenceBlock2D, ReferenceBlock3D,
                    RelevanceMask, apply_mask_2d, block_size_2d, get_gat_2d, get_hu_template_2d, get_input_image_2d,
                    get_input_image_3d, get_mean_noise_var_2d, get_reference_block_2d, get_reference_block_3d,
                    get_relevance_mask, normalize_image_2d)

__all__ = [
    "ABSDIFF", "BF16X3", "FP32", "NRMSE", "PEARSON", "SSIM", "Context", "NonBinaryMaskError", "SredError", "load",
    "BLOCK_METRICS", "GLOBAL_METRICS", "calculate_block_hu_map_2d", "calculate_block_repetition_map_2d",
    "calculate_global_hu_map_2d", "calculate_global_repetition_map_2d", "calculate_phase_correlation_map_2d",
    "get_hu_invariants_2d",
    "get_local_statistics_2d", "get_opencl_resources", "block_repetition_2d", "global_repetition_2d",
    "InputImage2D", "ReferenceBlock2D", "RelevanceMask", "apply_mask_2d", "block_size_2d", "get_hu_template_2d",
    "get_input_image_2d",
    "get_mean_noise_var_2d", "get_reference_block_2d", "get_relevance_mask", "normalize_image_2d",
    "GATMinimizer2D", "QuadTree", "get_gat_2d", "block_repetition_timelapse_2d", "global_repetition_3d",
    "block_repetition_3d", "InputImage3D", "ReferenceBlock3D", "get_input_image_3d", "get_reference_block_3d",
    "calculate_block_repetition_map_3d", "OcTree", "GATMinimizer3D", "variance_stabilising_transform_3d",
    "variance_stabilising_transform_2d", "quadtree_stabilise_2d", "octree_stabilise_3d",
]
