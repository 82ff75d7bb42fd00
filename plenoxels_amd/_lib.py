"""ctypes binding of libpxl (include/pxl.h).

torch is imported first on purpose: torch ships its own HIP runtime (libamdhip64.so.7)
and libpxl binds to that same already-loaded runtime through the shared soname, so the
streams and device pointers torch hands over are valid in the library.
"""
import ctypes
import os

import torch  # noqa: F401  (loads the HIP runtime libpxl binds to)

LIB_DIR = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib")
# PXL_LIBRARY names another build of the same library (e.g. the sanitizer build the ABI tests load)
LIB_PATH = os.environ.get("PXL_LIBRARY") or os.path.join(LIB_DIR, "libpxl.so")
PROBE_LIB_PATH = os.path.join(LIB_DIR, "libpxl_probe.so")

PXL_OK = 0
PXL_E_INVALID = -1
PXL_E_DTYPE = -2
PXL_E_SIZE = -3
PXL_E_HIP = -4
PXL_E_NO_DEVICE = -5

DTYPE_F32, DTYPE_BF16, DTYPE_F16 = 0, 1, 2
ABI_VERSION = 3  # PXL_ABI_VERSION of the include/pxl.h these signatures mirror
INTERP_NEAREST, INTERP_TRILINEAR = 0, 1
FLAG_EXACT_DIV = 1
REDUCE_OVERWRITE, REDUCE_ACCUMULATE, REDUCE_OVERWRITE_REACHABLE = 0, 1, 2
RAYS_CPU_NORM = 1
TRAIN_MAX_SAMPLES = 16384

c_void_p = ctypes.c_void_p


class Grid(ctypes.Structure):
    _fields_ = [("data", c_void_p), ("dims", ctypes.c_int64 * 3), ("channels", ctypes.c_int32),
                ("dtype", ctypes.c_int32), ("occupancy", c_void_p)]


class March(ctypes.Structure):
    _fields_ = [("n_samples", ctypes.c_int32), ("interp", ctypes.c_int32),
                ("flags", ctypes.c_int32), ("reserved", ctypes.c_int32),
                ("delta", ctypes.c_float), ("near", ctypes.c_float), ("t_stop", ctypes.c_float),
                ("pd", ctypes.c_float), ("inv_pd", ctypes.c_float),
                ("grid_min", ctypes.c_float * 3)]


class Rays(ctypes.Structure):
    _fields_ = [("n_rays", ctypes.c_int64), ("origins", c_void_p), ("dirs", c_void_p),
                ("cam_ids", c_void_p), ("uv", c_void_p), ("c2w", c_void_p),
                ("n_cams", ctypes.c_int64), ("camera_angle_x", ctypes.c_float),
                ("flags", ctypes.c_int32), ("imgs", c_void_p),
                ("img_h", ctypes.c_int64), ("img_w", ctypes.c_int64)]


class Loss(ctypes.Structure):
    _fields_ = [("targets", c_void_p), ("grad_scale", ctypes.c_float),
                ("beta_scale", ctypes.c_float), ("ray_sq_err", c_void_p), ("ray_beta", c_void_p)]


_GP = ctypes.POINTER(Grid)
_MP = ctypes.POINTER(March)
_RP = ctypes.POINTER(Rays)

# symbol -> (restype, argtypes); mirrors include/pxl.h
SIGNATURES = {
    "pxl_abi_version": (ctypes.c_int, []),
    "pxl_last_error": (ctypes.c_char_p, []),
    "pxl_aux_floats": (ctypes.c_int64, [ctypes.c_int32]),
    "pxl_render_fwd": (ctypes.c_int, [_GP, _MP, _RP, c_void_p, c_void_p, c_void_p]),
    "pxl_render_bwd": (ctypes.c_int, [_GP, _MP, _RP, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pxl_train_step": (ctypes.c_int, [_GP, _MP, _RP, ctypes.POINTER(Loss), c_void_p, c_void_p,
                                      c_void_p]),
    "pxl_train_step_blocks": (ctypes.c_int, [_GP, _MP, _RP, ctypes.POINTER(Loss), c_void_p, c_void_p,
                                             c_void_p]),
    "pxl_blocks_floats": (ctypes.c_int64, [_GP]),
    "pxl_scaled_sum": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_double, c_void_p, c_void_p,
                                      c_void_p]),
    "pxl_scaled_sum_work_bytes": (ctypes.c_int64, []),
    "pxl_scaled_sum_noalloc": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_double, c_void_p, c_void_p]),
    "pxl_scaled_sum_noalloc_scratch_bytes": (ctypes.c_int64, []),
    "pxl_blocks_reduce": (ctypes.c_int, [_GP, c_void_p, c_void_p, ctypes.c_int32, c_void_p, c_void_p]),
    "pxl_camera_rays": (ctypes.c_int, [_RP, c_void_p, c_void_p, c_void_p, c_void_p]),
    "pxl_blocks_clear": (ctypes.c_int, [_GP, c_void_p, c_void_p]),
    "pxl_occupancy_dilate": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_int64,
                                            ctypes.c_int64, c_void_p, c_void_p]),
    "pxl_pack_occupancy": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                          c_void_p, c_void_p]),
    "pxl_occupancy_words": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64]),
    "pxl_alpha_occupancy": (ctypes.c_int, [_GP, ctypes.c_int32, c_void_p, c_void_p]),
    "pxl_fill_f32": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_float, c_void_p]),
    "pxl_fill_reachable_f32": (ctypes.c_int, [_GP, ctypes.c_int32, c_void_p, ctypes.c_float, c_void_p]),
    "pxl_sample_points": (ctypes.c_int, [c_void_p, ctypes.c_int64, c_void_p, ctypes.c_int64,
                                         ctypes.c_int32, ctypes.c_float, ctypes.c_float, c_void_p,
                                         c_void_p]),
    "pxl_coord_min": (ctypes.c_int, [c_void_p, ctypes.c_int64, c_void_p, c_void_p]),
    "pxl_normalize_points": (ctypes.c_int, [c_void_p, ctypes.c_int64, c_void_p, ctypes.c_float,
                                            ctypes.c_float, ctypes.c_int32, c_void_p, c_void_p]),
    "pxl_in_bounds": (ctypes.c_int, [c_void_p, ctypes.c_int32, ctypes.c_int64,
                                     ctypes.POINTER(ctypes.c_int64), c_void_p, c_void_p]),
    "pxl_wrap_indices": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.POINTER(ctypes.c_int64),
                                        c_void_p]),
    "pxl_nearest_gather": (ctypes.c_int, [c_void_p, ctypes.c_int64, _GP, c_void_p, c_void_p, c_void_p]),
    "pxl_nearest_scatter": (ctypes.c_int, [c_void_p, ctypes.c_int64, _GP, c_void_p, c_void_p, c_void_p]),
    "pxl_corner_indices": (ctypes.c_int, [c_void_p, ctypes.c_int64, c_void_p, c_void_p]),
    "pxl_collect_cells": (ctypes.c_int, [c_void_p, ctypes.c_int64, _GP, c_void_p, c_void_p, c_void_p]),
    "pxl_collect_cells_bwd": (ctypes.c_int, [c_void_p, ctypes.c_int64, _GP, c_void_p, c_void_p,
                                             c_void_p]),
    "pxl_trilinear": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int64, _GP, c_void_p, c_void_p]),
    "pxl_trilinear_bwd": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int64, _GP, c_void_p, c_void_p,
                                         c_void_p]),
    "pxl_composite_fwd": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_int32, c_void_p, c_void_p]),
    "pxl_composite_bwd": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_int32, c_void_p, c_void_p,
                                         c_void_p]),
    "pxl_grid_coords": (ctypes.c_int, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, ctypes.c_float,
                                       c_void_p, c_void_p, c_void_p]),
    "pxl_avg_pool3d_workspace_floats": (ctypes.c_int64, [ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                                         ctypes.c_int32, ctypes.c_int32]),
    "pxl_avg_pool3d": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                      ctypes.c_int32, ctypes.c_int32, c_void_p, c_void_p, c_void_p]),
    "pxl_avg_pool3d_bwd": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64,
                                          ctypes.c_int32, ctypes.c_int32, c_void_p, ctypes.c_int32, c_void_p,
                                          c_void_p]),
    "pxl_tv_sumsq": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, c_void_p,
                                    c_void_p]),
    "pxl_tv_bwd": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, c_void_p,
                                  ctypes.c_float, c_void_p, c_void_p]),
    "pxl_adam_step": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int64,
                                     ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                     ctypes.c_int64, c_void_p, ctypes.c_int32, c_void_p]),
    "pxl_adam_tv_step": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, c_void_p, ctypes.c_float,
                                        ctypes.c_double, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                        ctypes.c_int64, c_void_p, ctypes.c_int32, ctypes.c_int32, c_void_p]),
    "pxl_count_cells_host": (ctypes.c_int, [_GP, _MP, c_void_p, c_void_p, ctypes.c_int64,
                                            c_void_p, c_void_p, ctypes.c_int32]),
}

# libpxl_probe (include/pxl_probe.h): measurement and test kernels, not loaded by the package
PROBE_SIGNATURES = {
    "pxl_probe_wave_ops": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "pxl_probe_memory_rate": (ctypes.c_int, [c_void_p, ctypes.c_int64, ctypes.c_int32, ctypes.c_int32,
                                             ctypes.c_int32, ctypes.c_int64, ctypes.c_int32, c_void_p, c_void_p]),
}


class PxlError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"libpxl error {code}: {msg}")
        self.code = code


_lib = None


def lib():
    """Load libpxl.so (built in-tree by __graft_entry__.build()); raise if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is not built; build it with "
                "`python -c 'import __graft_entry__ as g; g.build()'` from the repository root")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        if L.pxl_abi_version() != ABI_VERSION:
            raise ImportError(f"libpxl ABI version {L.pxl_abi_version()}, these bindings expect {ABI_VERSION}")
        _lib = L
    return _lib


_probe = None


def probe_lib():
    """Load libpxl_probe.so (measurement and test kernels, built beside libpxl.so)."""
    global _probe
    if _probe is None:
        if not os.path.exists(PROBE_LIB_PATH):
            raise ImportError(f"{PROBE_LIB_PATH} is not built; run __graft_entry__.build()")
        L = ctypes.CDLL(PROBE_LIB_PATH)
        for name, (res, args) in PROBE_SIGNATURES.items():
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _probe = L
    return _probe


def last_error():
    return lib().pxl_last_error().decode(errors="replace")


def check(rc):
    if rc != PXL_OK:
        raise PxlError(rc, last_error())
    return rc


def ptr(t):
    """Device pointer of a tensor (or None)."""
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream_handle(device=None):
    """hipStream_t of torch's current stream on `device`."""
    return ctypes.c_void_p(torch.cuda.current_stream(device).cuda_stream)


def dims_arg(dims):
    """int64[3] argument for the entry points that take grid dims."""
    return (ctypes.c_int64 * 3)(*[int(x) for x in dims])
