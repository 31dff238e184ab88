"""ctypes binding of ``libseiswave.so`` (declared in ``include/seiswave.h``).

The structures below mirror the header field for field.  The library is the
only compute path of this package: if it cannot be loaded, or no GPU is
visible, the solver classes raise instead of falling back to anything else.
"""
from __future__ import annotations

import ctypes as C
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libseiswave.so")

SW_MAX_ORDER = 8
SW_CLASS_CPML, SW_CLASS_TENSOR, SW_CLASS_SPML = 0, 1, 2
SW_SRC_NONE, SW_SRC_FORCE_VZ, SW_SRC_MOMENT, SW_SRC_EXPLOSIVE, SW_SRC_FORCE_VX = 0, 1, 2, 3, 4
SW_2D_VS, SW_2D_SV, SW_2D_RSG = 0, 1, 2
SW_RUN_FREE_SURFACE = 1
SW_MLP_NOALLOC_MAX_ROWS = 16384
FIELD_CODE = {"vx": 0, "vy": 1, "vz": 2, "sxx": 3, "syy": 4, "szz": 5, "sxy": 6, "sxz": 7, "syz": 8}

_fp = C.POINTER(C.c_float)
_u8p = C.POINTER(C.c_uint8)


class SwAxis(C.Structure):
    _fields_ = [(n, _fp) for n in ("a", "b", "k", "a_h", "b_h", "k_h", "d", "d_h")] + [
        ("in_", _u8p), ("in_h", _u8p), ("lo", C.c_int32), ("hi", C.c_int32)]


class SwSSG3D(C.Structure):
    _fields_ = (
        [(n, C.c_int32) for n in ("nx", "ny", "nz", "order", "cls")]
        + [(n, C.c_float) for n in ("dx", "dy", "dz", "dt")]
        + [("c", C.c_float * SW_MAX_ORDER)]
        + [(n, _fp) for n in ("rho", "lam", "mu", "vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")]
        + [("px", SwAxis), ("py", SwAxis), ("pz", SwAxis)]
        + [(n, _fp) for n in ("mem_x", "mem_y", "mem_z", "split")]
    )


class SwSSG2D(C.Structure):
    _fields_ = (
        [(n, C.c_int32) for n in ("nx", "nz", "order", "scheme")]
        + [(n, C.c_float) for n in ("dx", "dz", "dt")]
        + [("c", C.c_float * SW_MAX_ORDER)]
        + [(n, _fp) for n in ("rho", "lam", "mu", "vx", "vz", "sxx", "szz", "sxz")]
        + [("px", SwAxis), ("pz", SwAxis), ("mem", _fp)]
    )


class SwAdj2D(C.Structure):
    _fields_ = [(n, _fp) for n in ("vx", "vz", "sxx", "szz", "sxz", "mem", "tmp", "g_b1", "g_b4",
                                   "g_lamh", "g_muh", "g_muz", "g_src")]


class SwSource(C.Structure):
    _fields_ = [("kind", C.c_int32), ("ix", C.c_int32), ("iy", C.c_int32), ("iz", C.c_int32),
                ("mt", C.c_float * 6), ("amp", _fp), ("namp", C.c_int32)]


class SwReceivers(C.Structure):
    _fields_ = [("nrec", C.c_int32), ("ncomp", C.c_int32), ("comp", C.c_int32 * 9),
                ("idx", C.POINTER(C.c_int32)), ("data", _fp), ("nt", C.c_int32)]


class SwBand(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("ox", "oy", "oz", "ex", "ey", "ez")]


# Return types other than int.
RESTYPES = {"sw_ssg3d_split_cells": C.c_int64, "sw_ssg3d_split_map": C.c_int64}

# Every exported symbol of include/seiswave.h with its argument types.
_P3, _P2 = C.POINTER(SwSSG3D), C.POINTER(SwSSG2D)
_PS, _PR = C.POINTER(SwSource), C.POINTER(SwReceivers)
_PA = C.POINTER(SwAdj2D)
_PB = C.POINTER(SwBand)
SIGNATURES = {
    "sw_version": [],
    "sw_last_error": [C.c_char_p, C.c_size_t],
    "sw_device_init": [C.c_int],
    "sw_ssg3d_velocity": [_P3, C.c_void_p],
    "sw_ssg3d_stress": [_P3, C.c_void_p],
    "sw_ssg3d_velocity_spml": [_P3, C.c_void_p],
    "sw_ssg3d_stress_spml": [_P3, C.c_void_p],
    "sw_ssg3d_launch_plan": [_P3, C.POINTER(C.c_int32), C.c_int32],
    "sw_set_launch_plan": [C.c_int32],
    "sw_ssg3d_split_cells": [_P3],
    "sw_ssg3d_split_map": [_P3, C.c_int32, C.POINTER(C.c_int64), C.POINTER(C.c_int64), C.c_int64],
    "sw_ssg3d_inject": [_P3, _PS, C.c_int32, C.c_void_p],
    "sw_ssg3d_record": [_P3, _PR, C.c_int32, C.c_void_p],
    "sw_ssg3d_run": [_P3, _PS, _PR, C.c_int32, C.c_int32, C.c_void_p],
    "sw_ssg3d_run_ex": [_P3, _PS, _PR, C.c_int32, C.c_int32, C.c_int32, C.c_void_p],
    "sw_ssg2d_velocity": [_P2, C.c_void_p],
    "sw_ssg2d_stress": [_P2, C.c_void_p],
    "sw_ssg2d_inject": [_P2, _PS, C.c_int32, C.c_void_p],
    "sw_ssg2d_record": [_P2, _PR, C.c_int32, C.c_void_p],
    "sw_ssg2d_run": [_P2, _PS, _PR, C.c_int32, C.c_int32, C.c_void_p],
    "sw_ssg2d_free_surface_init": [_P2, C.c_void_p],
    "sw_ssg2d_free_surface": [_P2, C.c_void_p],
    "sw_ssg3d_free_surface_init": [_P3, _fp, _fp, C.c_void_p],
    "sw_ssg3d_free_surface": [_P3, C.c_void_p],
    "sw_ssg2d_adjoint_stress": [_P2, _PA, C.c_void_p],
    "sw_ssg2d_adjoint_velocity": [_P2, _PA, C.c_void_p],
    "sw_ssg2d_adjoint_inject": [_P2, _PA, _PR, C.c_int32, C.c_void_p],
    "sw_ssg2d_model_gradient": [_P2, _PA, _fp, _fp, _fp, _fp, _fp, C.c_void_p],
    "sw_ssg2d_adjoint_segment": [_P2, _PS, _PA, _PR, _fp, C.c_int32, C.c_int32, C.c_void_p],
    "sw_ssg2d_run_batch": [_P2, C.c_int32, C.c_int64, _PS, C.POINTER(C.c_int32), _PR, C.c_int64, C.c_int32,
                           C.c_int32, C.c_void_p],
    "sw_ssg2d_adjoint_segment_batch": [_P2, C.c_int32, C.c_int64, _PS, C.POINTER(C.c_int32), _PA, C.c_int64, _PR,
                                       C.c_int64, _fp, C.c_int32, C.c_int32, C.c_void_p],
    "sw_ssg2d_model_gradient_batch": [_P2, _PA, C.c_int32, C.c_int64, _fp, _fp, _fp, _fp, _fp, C.c_void_p],
    "sw_mlp_param_count": [],
    "sw_band_columns": [_PB],
    "sw_band_gather": [_P3, _PB, _fp, C.c_void_p],
    "sw_band_correct": [_P3, _PB, _fp, _fp, C.c_void_p],
    "sw_ssg3d_run_learned": [_P3, _PS, _PR, _PB, _fp, _fp, C.c_int32, C.c_int32, C.c_void_p],
    "sw_mlp_forward": [_fp, _fp, C.c_int32, _fp, _fp, _fp, C.c_void_p],
    "sw_mlp_loss_backward": [_fp] * 6 + [C.c_int32, _fp, _fp, _fp, C.c_void_p],
    "sw_mlp_loss_backward_noalloc": [_fp] * 6 + [C.c_int32, _fp, _fp, C.c_void_p],
    "sw_stream_probe": [C.POINTER(_fp), C.c_int32, C.POINTER(_fp), C.c_int32, C.c_int64, C.c_int32, C.c_void_p],
    "sw_mfma_probe": [_fp, C.c_int32, C.c_int32, C.c_void_p],
    "sw_adam": [_fp, _fp, _fp, _fp, C.c_int32, C.c_float, C.c_float, C.c_float, C.c_float, C.c_int32,
                C.c_void_p],
}

_lib = None


class SeisWaveError(RuntimeError):
    """A libseiswave call failed (invalid argument or HIP error)."""


def load(path: str | None = None):
    """Load libseiswave.so (once) and declare its signatures.

    Raises RuntimeError with the build command when the library is missing.
    """
    global _lib
    if _lib is not None and path is None:
        return _lib
    # torch first: the process must run one HIP runtime, and torch's is the one
    # that allocates the device memory this library is handed
    import torch  # noqa: F401
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(
            f"libseiswave.so not found at {p}; build it with "
            f"`make -C {os.path.join(_HERE, 'csrc')}` (or __graft_entry__.build())")
    lib = C.CDLL(p)
    for name, argtypes in SIGNATURES.items():
        fn = getattr(lib, name)
        fn.argtypes = argtypes
        fn.restype = RESTYPES.get(name, C.c_int)
    if path is None:
        _lib = lib
    return lib


def last_error(lib=None) -> str:
    lib = lib or load()
    buf = C.create_string_buffer(512)
    lib.sw_last_error(buf, 512)
    return buf.value.decode(errors="replace")


def check(rc: int, what: str) -> None:
    if rc != 0:
        raise SeisWaveError(f"{what}: {last_error()} (code {rc})")


def fptr(t) -> "C._Pointer":
    """float* of a contiguous float32 torch tensor (device or host)."""
    return C.cast(C.c_void_p(t.data_ptr()), _fp)


def u8ptr(t):
    return C.cast(C.c_void_p(t.data_ptr()), _u8p)


def i32ptr(t):
    return C.cast(C.c_void_p(t.data_ptr()), C.POINTER(C.c_int32))


def launch_plan(g) -> list:
    """CPML axis masks of the launches each 3-D sweep makes for descriptor
    ``g`` (one entry per launch; bit 0 x, 1 y, 2 z).  Host only."""
    lib = load()
    pm = (C.c_int32 * 8)()
    n = lib.sw_ssg3d_launch_plan(C.byref(g), pm, 8)
    if n < 0:
        check(n, "sw_ssg3d_launch_plan")
    return list(pm[:n])


def split_map(g, sweep=1):
    """(cell, slot) int64 arrays of every cell whose split values the
    split-field sweep ``sweep`` (0 velocity, 1 stress) updates for descriptor
    ``g``: cell = (k*ny + j)*nx + i, slot its index in the shell storage.
    Host only."""
    import numpy as np
    lib = load()
    n = lib.sw_ssg3d_split_map(C.byref(g), sweep, None, None, 0)
    if n < 0:
        check(int(n), "sw_ssg3d_split_map")
    cell = np.empty(n, np.int64)
    slot = np.empty(n, np.int64)
    p64 = C.POINTER(C.c_int64)
    m = lib.sw_ssg3d_split_map(C.byref(g), sweep, cell.ctypes.data_as(p64), slot.ctypes.data_as(p64), n)
    if m != n:
        raise SeisWaveError(f"sw_ssg3d_split_map returned {m} pairs, expected {n}")
    return cell, slot
