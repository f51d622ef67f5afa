"""ctypes binding of libsred_hip.so (the C ABI declared in include/sred.h).

The library is loaded from this package directory only.  There is no
fallback: if the HIP library is missing, importing the device entry points
raises, so nothing can silently run on the CPU.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libsred_hip.so")

SRED_OK = 0
SRED_ERR_ARG = -1
SRED_ERR_HIP = -2
SRED_ERR_NONBINARY_MASK = -3
SRED_ERR_UNKNOWN_METRIC = -4
SRED_ERR_NONFINITE = -5
SRED_ERR_NO_DEVICE = -6

PEARSON, SSIM, NRMSE, ABSDIFF = 0, 1, 2, 3
FP32, BF16X3 = 0, 1
OPT_CHUNK_COLS, OPT_SYMMETRIC, OPT_AUTOTUNE, OPT_BATCH_PANEL, OPT_BATCH_MEMORY_MB = 1, 2, 3, 4, 5
OPT_FIT_TIMING = 6
OPT_PARTIAL_MEMORY_MB = 7

# every symbol include/sred.h declares
EXPORTS = [
    "sred_version", "sred_last_error", "sred_ctx_create", "sred_ctx_destroy", "sred_ctx_set_stream",
    "sred_ctx_synchronize", "sred_ctx_set_chunk_cols", "sred_ctx_set_option", "sred_metric_from_name", "sred_block_size_2d",
    "sred_normalize_input_2d", "sred_reference_block_2d", "sred_noise_var_2d", "sred_noise_var_2d_dev", "sred_relevance_mask_2d",
    "sred_apply_mask_2d", "sred_normalize_map_2d", "sred_local_stats_2d", "sred_global_accum_2d",
    "sred_global_repetition_2d", "sred_block_repetition_2d", "sred_global_map_2d_dev", "sred_block_map_2d_dev",
    "sred_block_map_2d_dev_noalloc",
    "sred_ctx_last_contraction", "sred_global_plan_2d", "sred_global_pearson_adjoint_2d",
    "sred_global_pearson_adjoint_2d_dev", "sred_block_hu_2d", "sred_block_hu_2d_dev", "sred_hu_prepare_template_2d",
    "sred_global_batch_2d", "sred_ctx_tuned_panel", "sred_ctx_tuning_ms", "sred_hu_invariants_2d", "sred_global_hu_2d",
    "sred_global_hu_2d_dev", "sred_phase_correlation_2d", "sred_phase_correlation_2d_dev", "sred_ctx_last_skipped",
    "sred_gat_2d", "sred_gat_2d_dev", "sred_gat_objective_2d", "sred_gat_minimize_2d", "sred_gat_minimize_2d_dev",
    "sred_quadtree_gat_2d", "sred_quadtree_gat_2d_dev", "sred_gat_tree_2d", "sred_gat_tree_2d_dev", "sred_f_quantile",
    "sred_simple_regression", "sred_block_timelapse_2d", "sred_block_timelapse_2d_dev", "sred_tool_prep_2d",
    "sred_tool_prep_2d_dev", "sred_relevance_mask_tool_2d", "sred_relevance_mask_tool_2d_dev",
    "sred_optimise_relevance_mask_2d", "sred_optimise_relevance_mask_2d_dev", "sred_block_size_3d",
    "sred_local_stats_3d", "sred_global_repetition_3d", "sred_global_repetition_3d_dev",
    "sred_reference_block_3d", "sred_block_repetition_3d", "sred_block_repetition_3d_dev",
    "sred_octree_gat_3d", "sred_octree_gat_3d_dev", "sred_gat_tree_3d", "sred_gat_tree_3d_dev",
    "sred_gat_objective_3d", "sred_gat_minimize_3d", "sred_gat_minimize_3d_dev",
    "sred_relevance_mask_stack_2d", "sred_relevance_mask_stack_2d_dev",
]


class SredError(RuntimeError):
    """A non-zero status from the C ABI (message from sred_last_error)."""

    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class NonBinaryMaskError(SredError, ValueError):
    """Utils.normalizeImage2D's IllegalArgumentException (non-binary mask)."""


_lib = None
_F = ctypes.POINTER(ctypes.c_float)


class GlobalProblem(ctypes.Structure):
    """sred_global_problem: one image of a batched global call (device pointers)."""

    _fields_ = [("d_img", ctypes.c_void_p), ("w", ctypes.c_int), ("h", ctypes.c_int), ("bw", ctypes.c_int),
                ("bh", ctypes.c_int), ("d_out", ctypes.c_void_p)]
_VP = ctypes.c_void_p
_I = ctypes.c_int
_FL = ctypes.c_float
_I64 = ctypes.c_int64
_D = ctypes.c_double


def load():
    """Load libsred_hip.so, raising if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} not built: run `python -m sred_amd.build` (hipcc, gfx950)")
    # One HIP runtime per process: PyTorch wheels bundle their own HIP/HSA
    # runtime, loaded by the unversioned name libamdhip64.so.  Loading torch
    # first lets this library bind to that runtime through its soname
    # (libamdhip64.so.7); loading it first would put a second runtime into the
    # process and torch's device initialisation would then find no GPU.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    L = ctypes.CDLL(LIB_PATH)
    sig = {
        "sred_version": ([], _I),
        "sred_last_error": ([], ctypes.c_char_p),
        "sred_ctx_create": ([_I, ctypes.POINTER(_VP)], _I),
        "sred_ctx_destroy": ([_VP], _I),
        "sred_ctx_set_stream": ([_VP, _VP], _I),
        "sred_ctx_synchronize": ([_VP], _I),
        "sred_ctx_set_chunk_cols": ([_VP, _I], _I),
        "sred_ctx_set_option": ([_VP, _I, _I], _I),
        "sred_metric_from_name": ([ctypes.c_char_p, _I], _I),
        "sred_block_size_2d": ([_I, _I], _I),
        "sred_normalize_input_2d": ([_F, _I64], _I),
        "sred_reference_block_2d": ([_F, _I, _I, _F, _F, _F], _I),
        "sred_noise_var_2d": ([_F, _I, _I, _F], _I),
        "sred_noise_var_2d_dev": ([_VP, _VP, _I, _I, _F], _I),
        "sred_relevance_mask_2d": ([_F, _I, _I, _I, _I, _FL, _FL, _F, _F, _F], _I),
        "sred_apply_mask_2d": ([_F, _I, _I, _F], _I),
        "sred_normalize_map_2d": ([_F, _I, _I, _I, _I, _F], _I),
        "sred_local_stats_2d": ([_VP, _F, _I, _I, _I, _I, _F, _F], _I),
        "sred_global_accum_2d": ([_VP, _I, _I, _F, _F, _F, _I, _I, _I, _I, _FL, _FL, _F, _F], _I),
        "sred_global_repetition_2d": ([_VP, _I, _I, _F, _F, _F, _I, _I, _I, _I, _F, _FL, _FL, _FL, _FL, _I, _F],
                                      _I),
        "sred_block_repetition_2d": ([_VP, _I, _F, _I, _I, _F, _I, _I, _FL, _FL, _FL, _I, _F], _I),
        "sred_global_map_2d_dev": ([_VP, _I, _I, _VP, _I, _I, _I, _I, _FL, _I, _VP, _VP, _VP, _VP, _VP, _VP], _I),
        "sred_block_map_2d_dev": ([_VP, _I, _VP, _I, _I, _F, _I, _I, _FL, _FL, _FL, _I, _VP], _I),
        "sred_block_map_2d_dev_noalloc": ([_VP, _I, _VP, _I, _I, _F, _I, _I, _FL, _FL, _FL, _I, _VP], _I),
        "sred_ctx_last_contraction": ([_VP, _F, ctypes.POINTER(_I), ctypes.POINTER(_I64)], _I),
        "sred_ctx_last_skipped": ([_VP, ctypes.POINTER(_I64)], _I),
        "sred_gat_2d": ([_VP, _F, _I64, _D, _D, _D], _I),
        "sred_gat_2d_dev": ([_VP, _VP, _I64, _D, _D, _D], _I),
        "sred_gat_objective_2d": ([_VP, _F, _I, _I, _D, _D, _D, ctypes.POINTER(_D)], _I),
        "sred_gat_minimize_2d": ([_VP, _F, _I, _I, _D, _D, _D, _I, ctypes.POINTER(_D), ctypes.POINTER(_D)], _I),
        "sred_gat_minimize_2d_dev": ([_VP, _VP, _I, _I, _D, _D, _D, _I, ctypes.POINTER(_D), ctypes.POINTER(_D)],
                                     _I),
        "sred_quadtree_gat_2d": ([_VP, _F, _I, _I, _I, _FL, _I, _FL, ctypes.POINTER(_D), ctypes.POINTER(_I), _I,
                                  _VP, _F, _F], _I),
        "sred_quadtree_gat_2d_dev": ([_VP, _VP, _I, _I, _I, _FL, _I, _FL, ctypes.POINTER(_D), ctypes.POINTER(_I),
                                      _I, _VP, _F, _F], _I),
        "sred_gat_tree_2d": ([_VP, _F, _I64, _D, _D], _I),
        "sred_block_timelapse_2d": ([_VP, _I, _F, _I, _I, _I, _F, _I, _I, _FL, _FL, _FL, _I, _F], _I),
        "sred_tool_prep_2d": ([_VP, _F, _I, _I, _I, ctypes.POINTER(_D)], _I),
        "sred_block_size_3d": ([_I, _I, _I], _I),
        "sred_reference_block_3d": ([_F, _I, _I, _I, _F, _F, _F], _I),
        "sred_octree_gat_3d": ([_VP, _F, _I, _I, _I, _I, _FL, _I, _FL, _VP, _VP, _I, _VP, _VP, _VP], _I),
        "sred_octree_gat_3d_dev": ([_VP, _VP, _I, _I, _I, _I, _FL, _I, _FL, _VP, _VP, _I, _VP, _VP, _VP], _I),
        "sred_gat_tree_3d": ([_VP, _F, _I64, _FL, _FL], _I),
        "sred_gat_objective_3d": ([_VP, _F, _I, _I, _I, _D, _D, _D, _VP], _I),
        "sred_relevance_mask_stack_2d": ([_VP, _F, _I, _I, _I, _I, _FL, _FL, _FL, _I, _VP, _VP, _VP], _I),
        "sred_relevance_mask_stack_2d_dev": ([_VP, _VP, _I, _I, _I, _I, _FL, _FL, _FL, _I, _VP, _VP, _VP], _I),
        "sred_gat_minimize_3d": ([_VP, _F, _I, _I, _I, _D, _D, _D, _I, _VP, _VP], _I),
        "sred_gat_minimize_3d_dev": ([_VP, _VP, _I, _I, _I, _D, _D, _D, _I, _VP, _VP], _I),
        "sred_gat_tree_3d_dev": ([_VP, _VP, _I64, _FL, _FL], _I),
        "sred_block_repetition_3d": ([_VP, _I, _F, _I, _I, _I, _F, _I, _I, _I, _FL, _FL, _FL, _I, _F], _I),
        "sred_block_repetition_3d_dev": ([_VP, _I, _VP, _I, _I, _I, _F, _I, _I, _I, _FL, _FL, _FL, _I, _VP, _VP,
                                          _VP, _VP], _I),
        "sred_local_stats_3d": ([_VP, _F, _I, _I, _I, _I, _I, _I, _F, _F], _I),
        "sred_global_repetition_3d": ([_VP, _I, _I, _F, _I, _I, _I, _I, _I, _I, _FL, _I, _F], _I),
        "sred_global_repetition_3d_dev": ([_VP, _I, _I, _VP, _I, _I, _I, _I, _I, _I, _FL, _I, _VP, _VP, _VP, _VP,
                                           _VP, _VP], _I),
        "sred_tool_prep_2d_dev": ([_VP, _VP, _I, _I, _I, ctypes.POINTER(_D)], _I),
        "sred_relevance_mask_tool_2d": ([_VP, _F, _I, _I, _I, _I, _FL, _F, _F, _F], _I),
        "sred_relevance_mask_tool_2d_dev": ([_VP, _VP, _I, _I, _I, _I, _FL, _VP, _F, _F], _I),
        "sred_optimise_relevance_mask_2d": ([_VP, _F, _I, _I, _I, _I, _FL, _I64, _F, _F], _I),
        "sred_optimise_relevance_mask_2d_dev": ([_VP, _VP, _I, _I, _I, _I, _FL, _I64, _F, _F], _I),
        "sred_block_timelapse_2d_dev": ([_VP, _I, _VP, _I, _I, _I, _F, _I, _I, _FL, _FL, _FL, _I, _VP], _I),
        "sred_gat_tree_2d_dev": ([_VP, _VP, _I64, _D, _D], _I),
        "sred_f_quantile": ([_D, _D, _D, ctypes.POINTER(_D)], _I),
        "sred_simple_regression": ([ctypes.POINTER(_D), ctypes.POINTER(_D), _I64, ctypes.POINTER(_D),
                                    ctypes.POINTER(_D)], _I),
        "sred_global_plan_2d": ([_I, _I, _I, _I, _I, _I, ctypes.POINTER(_I64)], _I),
        "sred_global_pearson_adjoint_2d": ([_VP, _F, _I, _I, _I, _I, _F, _F], _I),
        "sred_global_pearson_adjoint_2d_dev": ([_VP, _VP, _I, _I, _I, _I, _VP, _VP, _VP], _I),
        "sred_block_hu_2d": ([_VP, _F, _I, _I, _F, _I, _I, _F], _I),
        "sred_block_hu_2d_dev": ([_VP, _VP, _I, _I, _F, _I, _I, _VP], _I),
        "sred_hu_prepare_template_2d": ([_F, _I, _I, _F], _I),
        "sred_global_batch_2d": ([_VP, _I, _I, ctypes.POINTER(GlobalProblem), _FL, _I], _I),
        "sred_ctx_tuned_panel": ([_VP, _I, _I, ctypes.POINTER(_I)], _I),
        "sred_ctx_tuning_ms": ([_VP, _I, _I, _F], _I),
        "sred_hu_invariants_2d": ([_VP, _F, _F, _I, _I, _I, _I, _F], _I),
        "sred_global_hu_2d": ([_VP, _F, _F, _F, _I, _I, _I, _I, _VP, _I64, _F], _I),
        "sred_global_hu_2d_dev": ([_VP, _VP, _VP, _VP, _I, _I, _I, _I, _VP, _I64, _VP], _I),
        "sred_phase_correlation_2d": ([_VP, _F, _F, _F, _I, _I, _I, _I, _VP, _I64, _F], _I),
        "sred_phase_correlation_2d_dev": ([_VP, _VP, _VP, _VP, _I, _I, _I, _I, _VP, _I64, _VP], _I),
    }
    for name, (args, res) in sig.items():
        fn = getattr(L, name)
        fn.argtypes = args
        fn.restype = res
    _lib = L
    return L


def check(rc):
    """Raise SredError for a negative status; return rc otherwise."""
    if rc < 0:
        msg = load().sred_last_error().decode(errors="replace")
        if rc == SRED_ERR_NONBINARY_MASK:
            raise NonBinaryMaskError(rc, msg)
        raise SredError(rc, msg)
    return rc


def global_plan(n_valid, block_pixels, metric, precision, symmetric=1, chunk_cols=0):
    """sred_global_plan_2d as a dict (host only, no GPU)."""
    out = (_I64 * 8)()
    check(load().sred_global_plan_2d(n_valid, block_pixels, metric, precision, symmetric, chunk_cols, out))
    keys = ["nv_pad", "rows", "cols_per_chunk", "nchunks", "grid", "partial_elems", "pairs", "flags"]
    return dict(zip(keys, list(out)))


def fptr(a):
    """float* of a C-contiguous float32 numpy array (or None)."""
    if a is None:
        return None
    assert a.dtype == np.float32 and a.flags["C_CONTIGUOUS"], "expected C-contiguous float32"
    return a.ctypes.data_as(_F)


def f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


class Context:
    """A device context (sred_ctx): CLUtils.OpenCLResources' counterpart, reusable."""

    def __init__(self, device=0):
        self._lib = load()
        h = _VP()
        check(self._lib.sred_ctx_create(device, ctypes.byref(h)))
        self.handle = h
        self.device = device

    def close(self):
        if self.handle:
            self._lib.sred_ctx_destroy(self.handle)
            self.handle = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_stream(self, stream_ptr):
        check(self._lib.sred_ctx_set_stream(self.handle, _VP(stream_ptr)))

    def synchronize(self):
        check(self._lib.sred_ctx_synchronize(self.handle))

    def set_chunk_cols(self, cols):
        check(self._lib.sred_ctx_set_chunk_cols(self.handle, int(cols)))

    def set_option(self, option, value):
        check(self._lib.sred_ctx_set_option(self.handle, int(option), int(value)))

    def tuned_panel(self, metric, k_steps):
        """Panel size the autotuner chose for (metric, ceil(K/32)), 0 if none yet."""
        p = ctypes.c_int()
        check(self._lib.sred_ctx_tuned_panel(self.handle, int(metric), int(k_steps), ctypes.byref(p)))
        return p.value

    def tuning_ms(self, metric, k_steps):
        """The autotuner's timed launch per candidate panel (256, 512, 1024, 2048), ms; 0 if not timed."""
        ms = (ctypes.c_float * 4)()
        check(self._lib.sred_ctx_tuning_ms(self.handle, int(metric), int(k_steps), ms))
        return dict(zip((256, 512, 1024, 2048), list(ms)))

    def last_contraction(self):
        """(ms, launches, pairs_evaluated) of the last global contraction."""
        ms = ctypes.c_float()
        n = ctypes.c_int()
        p = ctypes.c_int64()
        check(self._lib.sred_ctx_last_contraction(self.handle, ctypes.byref(ms), ctypes.byref(n), ctypes.byref(p)))
        return ms.value, n.value, p.value

    def last_skipped(self):
        """Ordered block pairs the last global call skipped (each term below 2^-40)."""
        p = ctypes.c_int64()
        check(self._lib.sred_ctx_last_skipped(self.handle, ctypes.byref(p)))
        return p.value
