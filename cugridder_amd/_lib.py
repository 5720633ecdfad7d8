"""Loads libcurafft.so (built in-tree) and declares its C-ABI.

The library is the product: if it is missing this module raises instead of
falling back to anything else.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcurafft.so")
_lib = None

c_int, c_float, c_double = ctypes.c_int, ctypes.c_float, ctypes.c_double
c_void_p, c_char_p = ctypes.c_void_p, ctypes.c_char_p


def _nd(dtype):
    return np.ctypeslib.ndpointer(dtype, flags="C")


def _nd_or_none(dtype):
    """ndpointer that also accepts None (NULL)."""
    base = _nd(dtype)

    class _P(base):
        @classmethod
        def from_param(cls, obj):
            if obj is None:
                return None
            return base.from_param(obj)

    return _P


class BatchItem(ctypes.Structure):
    """cg_batch_item of include/curafft.h."""
    _fields_ = [("nrow", ctypes.c_int), ("nxdirty", ctypes.c_int), ("nydirty", ctypes.c_int),
                ("fov", ctypes.c_double), ("freq", ctypes.c_double), ("d_uvw", ctypes.c_void_p),
                ("d_vis", ctypes.c_void_p), ("d_wgt", ctypes.c_void_p), ("d_dirty", ctypes.c_void_p)]


BATCH_AUTOTUNE = 1
BATCH_GRAPH = 2


def lib():
    """The loaded libcurafft.so with argtypes declared."""
    global _lib
    if _lib is not None:
        return _lib
    # One HIP runtime per process: torch ships its own libamdhip64 with the
    # same SONAME (libamdhip64.so.7); loading torch first makes this library
    # bind to that runtime, so device pointers are shared with torch tensors.
    try:
        import torch  # noqa: F401
    except ImportError:
        pass
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            "libcurafft.so is not built (expected at %s); run "
            "`python -c 'import __graft_entry__ as g; g.build()'`" % LIB_PATH)
    L = ctypes.CDLL(LIB_PATH)
    f64, f32 = _nd(np.float64), _nd(np.float32)
    c128, c64 = _nd(np.complex128), _nd(np.complex64)
    f64n, f32n = _nd_or_none(np.float64), _nd_or_none(np.float32)
    # reference signatures (python/curagridder/cursl.py:29-72)
    L.ms2dirty_1.argtypes = [c_int, c_int, c_int, c_double, c_double, f64, c128, c128,
                             c_double, c_double, c_int]
    L.ms2dirty_2.argtypes = [c_int, c_int, c_int, c_double, c_double, f64, c128, f64n, c128,
                             c_double, c_double, c_int]
    L.dirty2ms_1.argtypes = [c_int, c_int, c_int, c_double, c_double, f64, c128, c128,
                             c_double, c_double, c_int]
    L.dirty2ms_2.argtypes = [c_int, c_int, c_int, c_double, c_double, f64, c128, f64n, c128,
                             c_double, c_double, c_int]
    L.ms2dirtyf_1.argtypes = [c_int, c_int, c_int, c_float, c_float, f32, c64, c64, c_float,
                              c_float, c_int]
    L.ms2dirtyf_2.argtypes = [c_int, c_int, c_int, c_float, c_float, f32, c64, f32n, c64,
                              c_float, c_float, c_int]
    L.dirty2msf_1.argtypes = [c_int, c_int, c_int, c_float, c_float, f32, c64, c64, c_float,
                              c_float, c_int]
    L.dirty2msf_2.argtypes = [c_int, c_int, c_int, c_float, c_float, f32, c64, f32n, c64,
                              c_float, c_float, c_int]
    L.ms2dirty_mixed.argtypes = [c_int, c_int, c_int, c_double, c_double, f64, c64, f32n, c64,
                                 c_double, c_double, c_int]
    L.dirty2ms_mixed.argtypes = [c_int, c_int, c_int, c_double, c_double, f64, c64, f32n, c64,
                                 c_double, c_double, c_int]
    # several channels: freq[nchan], vis/wgt [nrow][nchan]
    L.ms2dirty_mc_2.argtypes = [c_int, c_int, c_int, c_int, c_double, f64, f64, c128, f64n, c128,
                                c_double, c_double, c_int]
    L.dirty2ms_mc_2.argtypes = [c_int, c_int, c_int, c_int, c_double, f64, f64, c128, f64n, c128,
                                c_double, c_double, c_int]
    L.ms2dirty_mc_mixed.argtypes = [c_int, c_int, c_int, c_int, c_double, f64, f64, c64, f32n,
                                    c64, c_double, c_double, c_int]
    L.dirty2ms_mc_mixed.argtypes = [c_int, c_int, c_int, c_int, c_double, f64, f64, c64, f32n,
                                    c64, c_double, c_double, c_int]
    L.nufft2d1.argtypes = [c_int, f64, f64, c128, c_int, c_int, c128, c_double, c_double]
    L.nufft2d2.argtypes = [c_int, f64, f64, c128, c_int, c_int, c128, c_double, c_double]
    L.nufft2d1f.argtypes = [c_int, f32, f32, c64, c_int, c_int, c64, c_float, c_float]
    L.nufft2d2f.argtypes = [c_int, f32, f32, c64, c_int, c_int, c64, c_float, c_float]
    # separate pixel sizes and plain 2-D gridding (prec 0 mixed, 1 double,
    # 2 float); uvw / vis / wgt / dirty are raw host pointers
    L.ms2dirty_px.argtypes = [c_int, c_int, c_int, c_int, c_int, c_double, c_double, f64,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_int,
                              c_int]
    L.dirty2ms_px.argtypes = [c_int, c_int, c_int, c_int, c_int, c_double, c_double, f64,
                              c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double, c_int,
                              c_int]
    for name in ("ms2dirty_px", "dirty2ms_px", "ms2dirty_1", "ms2dirty_2", "dirty2ms_1",
                 "dirty2ms_2", "ms2dirtyf_1",
                 "ms2dirtyf_2", "dirty2msf_1", "dirty2msf_2", "ms2dirty_mixed",
                 "dirty2ms_mixed", "ms2dirty_mc_2", "dirty2ms_mc_2", "ms2dirty_mc_mixed",
                 "dirty2ms_mc_mixed", "nufft2d1", "nufft2d2", "nufft2d1f", "nufft2d2f"):
        getattr(L, name).restype = c_int
    # contexts / device pointers
    L.cg_ctx_create.argtypes = [ctypes.POINTER(c_void_p), c_void_p]
    L.cg_ctx_destroy.argtypes = [c_void_p]
    L.cg_ctx_set_option.argtypes = [c_void_p, c_char_p, c_int]
    L.cg_ctx_stage_times.argtypes = [c_void_p, f64, ctypes.POINTER(ctypes.c_longlong)]
    L.cg_ctx_reset_times.argtypes = [c_void_p]
    L.cg_ctx_last_plan.argtypes = [c_void_p, f64]
    L.cg_ctx_last_band.argtypes = [c_void_p, np.ctypeslib.ndpointer(np.int32, flags="C")]
    L.cg_ctx_last_slab.argtypes = [c_void_p, np.ctypeslib.ndpointer(np.int32, flags="C")]
    L.cg_ms2dirty_dev.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_double, c_double,
                                  c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                  c_double, c_double, c_int]
    L.cg_dirty2ms_dev.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_double, c_double,
                                  c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                  c_double, c_double, c_int]
    L.cg_ms2dirty_mc_dev.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int, c_double, f64,
                                     c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                     c_double, c_double, c_int]
    L.cg_dirty2ms_mc_dev.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int, c_double, f64,
                                     c_void_p, c_int, c_void_p, c_void_p, c_int, c_void_p,
                                     c_double, c_double, c_int]
    L.cg_ms2dirty_px_dev.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int, c_double,
                                     c_double, f64, c_void_p, c_int, c_void_p, c_void_p, c_int,
                                     c_void_p, c_double, c_double, c_int, c_int]
    L.cg_dirty2ms_px_dev.argtypes = [c_void_p, c_int, c_int, c_int, c_int, c_int, c_double,
                                     c_double, f64, c_void_p, c_int, c_void_p, c_void_p, c_int,
                                     c_void_p, c_double, c_double, c_int, c_int]
    L.cg_nufft2d_dev.argtypes = [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p, c_void_p,
                                 c_int, c_int, c_void_p, c_double, c_double]
    L.cg_nufft2d_noalloc.argtypes = [c_void_p, c_int, c_int, c_int, c_void_p, c_void_p,
                                     c_void_p, c_int, c_int, c_void_p, c_double, c_double]
    L.cg_batch_create.argtypes = [ctypes.POINTER(c_void_p), c_void_p, c_int, c_int,
                                  ctypes.POINTER(BatchItem), c_double, c_double, c_int, c_int]
    L.cg_batch_run.argtypes = [c_void_p]
    L.cg_batch_info.argtypes = [c_void_p, f64]
    L.cg_batch_destroy.argtypes = [c_void_p]
    L.cg_plan_query.argtypes = [c_int, c_int, c_int, c_double, c_double, c_double, c_double,
                                c_double, c_double, c_int, c_int, f64]
    L.cg_ctx_grid.argtypes = [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(ctypes.c_longlong)]
    L.cg_ctx_sorted.argtypes = [c_void_p, ctypes.POINTER(c_void_p), ctypes.POINTER(c_void_p),
                                ctypes.POINTER(c_void_p), ctypes.POINTER(ctypes.c_longlong)]
    L.cg_fft2d_dev.argtypes = [c_void_p, c_int, c_void_p, c_int, c_int, c_int, c_int]
    L.cg_sort_pairs_dev.argtypes = [c_void_p, c_void_p, c_void_p, c_int, c_int]
    L.cg_memcpy_dev.argtypes = [c_void_p, c_void_p, c_void_p, ctypes.c_size_t]
    L.cg_fft_col_split.argtypes = [c_int, ctypes.POINTER(c_int), ctypes.POINTER(c_int)]
    L.cg_tap_poly.argtypes = [c_int, c_double, c_double, f32n, c_int, c_int, f32]
    # learned kernel
    L.cg_learn_grad.argtypes = [c_int, f32, f64, f64, c_int, f32, f64]
    L.cg_learn_train.argtypes = [c_int, c_int, f64, f64, c_int, c_int, c_int, c_double, f32, f64n]
    L.cg_learn_eval.argtypes = [f32, f32, c_int, f32]
    L.cg_learn_tabulate.argtypes = [c_int, c_int, f32, f32, f64]
    L.cg_ctx_set_kernel.argtypes = [c_void_p, c_int, c_int, f32n, c_double]
    for name in ("cg_ctx_create", "cg_ctx_destroy", "cg_ctx_set_option", "cg_ctx_stage_times",
                 "cg_ctx_reset_times", "cg_ctx_last_plan", "cg_ctx_last_band", "cg_ctx_last_slab", "cg_ms2dirty_dev", "cg_dirty2ms_dev",
                 "cg_ms2dirty_mc_dev", "cg_dirty2ms_mc_dev", "cg_ms2dirty_px_dev",
                 "cg_dirty2ms_px_dev", "cg_batch_create", "cg_batch_run",
                 "cg_batch_info", "cg_batch_destroy",
                 "cg_nufft2d_dev", "cg_nufft2d_noalloc", "cg_plan_query", "cg_ctx_grid", "cg_ctx_sorted",
                 "cg_fft2d_dev", "cg_sort_pairs_dev", "cg_memcpy_dev", "cg_fft_col_split", "cg_tap_poly",
                 "cg_version", "cg_learn_nparams", "cg_learn_grad", "cg_learn_train",
                 "cg_learn_eval", "cg_learn_tabulate", "cg_ctx_set_kernel"):
        getattr(L, name).restype = c_int
    _lib = L
    return L


ERRORS = {1: "epsilon clamped", 2: "invalid argument", 3: "unsupported size",
          4: "HIP runtime error", 5: "device allocation failed"}


def check(ier, what):
    """Raise on fatal codes (>= 2); code 1 (eps clamped) is a warning."""
    if ier >= 2:
        raise RuntimeError("%s failed: ier=%d (%s)" % (what, ier, ERRORS.get(ier, "?")))
    return ier
