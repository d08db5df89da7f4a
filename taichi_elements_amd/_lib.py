"""ctypes binding of libmpm_hip.so (declared in include/mpm_abi.h).

There is no fallback: if the HIP library is missing or cannot be loaded,
importing the solver fails loudly.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmpm_hip.so")

MPM_OK, MPM_EINVAL, MPM_EHIP, MPM_EDOMAIN, MPM_ENOMEM = 0, -1, -2, -3, -4
FIELD_X, FIELD_V, FIELD_C, FIELD_F, FIELD_JP, FIELD_MATERIAL, FIELD_COLOR = range(7)


class MpmConfig(ctypes.Structure):
    _fields_ = [("dim", ctypes.c_int32), ("res", ctypes.c_int32 * 3),
                ("padding", ctypes.c_int32), ("device", ctypes.c_int32),
                ("reorder_interval", ctypes.c_int32), ("chunk", ctypes.c_int32),
                ("dx", ctypes.c_double), ("inv_dx", ctypes.c_double),
                ("p_vol", ctypes.c_double), ("p_mass", ctypes.c_double),
                ("mu_0", ctypes.c_double), ("lambda_0", ctypes.c_double),
                ("alpha", ctypes.c_double), ("water_density", ctypes.c_double)]


# (name, restype, argtypes) of every exported entry point
_P = ctypes.c_void_p
_D3 = ctypes.POINTER(ctypes.c_double)
_I32P = ctypes.POINTER(ctypes.c_int32)
_I64P = ctypes.POINTER(ctypes.c_int64)
_FP = ctypes.POINTER(ctypes.c_float)
SIGNATURES = [
    ("mpm_create", ctypes.c_int, [ctypes.POINTER(MpmConfig), ctypes.POINTER(_P)]),
    ("mpm_create_batch", ctypes.c_int, [ctypes.POINTER(MpmConfig), ctypes.c_int32,
                                        ctypes.POINTER(_P)]),
    ("mpm_set_active_scene", ctypes.c_int, [_P, ctypes.c_int32]),
    ("mpm_substeps_batch", ctypes.c_int, [_P, ctypes.c_int32, _FP]),
    ("mpm_destroy", ctypes.c_int, [_P]),
    ("mpm_last_error", ctypes.c_char_p, []),
    ("mpm_set_gravity", ctypes.c_int, [_P, _D3]),
    ("mpm_add_bounding_box", ctypes.c_int, [_P, ctypes.c_int32]),
    ("mpm_add_surface_collider", ctypes.c_int, [_P, _D3, _D3, ctypes.c_int32, ctypes.c_double]),
    ("mpm_add_sphere_collider", ctypes.c_int, [_P, _D3, ctypes.c_double, ctypes.c_int32]),
    ("mpm_clear_grid_postprocess", ctypes.c_int, [_P]),
    ("mpm_add_particles", ctypes.c_int, [_P, ctypes.c_int64, _FP, _FP, _I32P, _I32P]),
    ("mpm_num_particles", ctypes.c_int, [_P, _I64P]),
    ("mpm_substeps", ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_float, ctypes.c_double]),
    ("mpm_read_max_velocity", ctypes.c_int, [_P, ctypes.c_int32, _FP]),
    ("mpm_max_grid_velocity", ctypes.c_int, [_P, _FP]),
    ("mpm_substep_begin", ctypes.c_int, [_P, ctypes.c_float]),
    ("mpm_grid_node_count", ctypes.c_int, [_P, _I64P]),
    ("mpm_grid_gather", ctypes.c_int, [_P, _P, _P, _P, ctypes.c_int64]),
    ("mpm_grid_scatter", ctypes.c_int, [_P, _P, ctypes.c_int64]),
    ("mpm_substep_end", ctypes.c_int, [_P]),
    ("mpm_particle_info", ctypes.c_int, [_P, _FP, _FP, _I32P, _I32P]),
    ("mpm_get_state", ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_void_p]),
    ("mpm_set_state", ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_void_p]),
    ("mpm_read_grid", ctypes.c_int, [_P, _I32P, _I32P, _FP, _FP]),
    ("mpm_read_binning", ctypes.c_int, [_P, _I32P, _I32P, _I32P]),
    ("mpm_set_debug", ctypes.c_int, [_P, ctypes.c_int32]),
    ("mpm_sync", ctypes.c_int, [_P]),
    ("mpm_set_timing", ctypes.c_int, [_P, ctypes.c_int32]),
    ("mpm_kernel_times", ctypes.c_int, [_P, ctypes.c_char_p, ctypes.c_int32, _I32P,
                                        ctypes.POINTER(ctypes.c_double), ctypes.c_int32, _I32P]),
    ("mpm_last_counts", ctypes.c_int, [_P, _I64P]),
    ("mpm_host_svd", ctypes.c_int, [ctypes.c_int32, _FP, _FP, _FP, _FP]),
    ("mpm_set_g2p2g", ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_double, ctypes.c_int32]),
    ("mpm_set_graphs", ctypes.c_int, [_P, ctypes.c_int32]),
    ("mpm_set_tuning", ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    ("mpm_get_tuning", ctypes.c_int, [_P, _I32P, _I32P]),
    ("mpm_set_g2p_waves", ctypes.c_int, [_P, ctypes.c_int32]),
    ("mpm_get_g2p_waves", ctypes.c_int, [_P, _I32P]),
    ("mpm_snapshot", ctypes.c_int, [_P, ctypes.c_int32]),
    ("mpm_device_svd", ctypes.c_int, [ctypes.c_int32, ctypes.c_int32, ctypes.c_int64, _FP, _FP,
                                      _FP, _FP]),
    ("mpm_diff_setup", ctypes.c_int, [_P, ctypes.c_int64, _I32P]),
    ("mpm_diff_binning_bytes", ctypes.c_int, [_P, _I64P]),
    ("mpm_diff_substep", ctypes.c_int, [_P, ctypes.c_float, _P, _P, _P, ctypes.c_double, _P,
                                        _P]),
    ("mpm_diff_substep_backward", ctypes.c_int, [_P, ctypes.c_float, _P, _P, ctypes.c_double, _P,
                                                 _P, _P, _P, _P, _P]),
    ("mpm_diff_sync", ctypes.c_int, [_P, _P]),
    ("mpm_diff_substep_noalloc", ctypes.c_int, [_P, ctypes.c_float, _P, _P, _P,
                                                ctypes.c_double, _P]),
    ("mpm_quantize_particles", ctypes.c_int, [_P, _FP, ctypes.POINTER(ctypes.c_uint32),
                                              ctypes.POINTER(ctypes.c_uint8)]),
    ("mpm_set_storage", ctypes.c_int, [_P, ctypes.c_int32, ctypes.c_int32]),
    ("mpm_set_unbounded", ctypes.c_int, [_P, ctypes.c_int32]),
    ("mpm_particle_bytes", ctypes.c_int, [_P, _I32P]),
    ("mpm_voxelize", ctypes.c_int, [ctypes.c_int32, _I32P, ctypes.c_double, ctypes.c_int32,
                                    ctypes.c_int32, ctypes.c_int64, _D3, _I32P, _I32P]),
    ("mpm_add_mesh", ctypes.c_int, [_P, ctypes.c_int64, _D3, ctypes.c_int32, ctypes.c_int32,
                                    ctypes.c_int32, _FP, _FP, ctypes.c_int32, ctypes.c_uint32,
                                    _I64P]),
]

_lib = None


def lib():
    """Load libmpm_hip.so once; raises ImportError if it is missing."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with "
                              "`python -m taichi_elements_amd.build` (or __graft_entry__.build())")
        L = ctypes.CDLL(LIB_PATH)
        for name, res, args in SIGNATURES:
            fn = getattr(L, name)
            fn.restype = res
            fn.argtypes = args
        _lib = L
    return _lib


class MPMError(RuntimeError):
    pass


class DomainError(MPMError):
    """Particles left the simulated grid window."""


def check(rc):
    if rc == MPM_OK:
        return
    msg = lib().mpm_last_error().decode(errors="replace")
    if rc == MPM_EINVAL:
        raise ValueError(msg)
    if rc == MPM_EDOMAIN:
        raise DomainError(msg)
    if rc == MPM_ENOMEM:
        raise MemoryError(msg)
    raise MPMError(f"HIP error: {msg}")


def fptr(a):
    return a.ctypes.data_as(_FP) if a is not None else None


def iptr(a):
    return a.ctypes.data_as(_I32P) if a is not None else None


def dvec(values, n=3):
    arr = (ctypes.c_double * n)()
    for i, v in enumerate(values):
        arr[i] = float(v)
    return arr


def as_f32(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.float32).reshape(shape))


def as_i32(a, shape):
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32).reshape(shape))
