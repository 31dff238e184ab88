"""A stand-in for the part of Taichi that the reference's driver scripts use
outside the solver classes, so such a script runs without Taichi installed.

The drivers (test3d_homegenous_tensor.py:1-91 and the other test*.py scripts)
call ``ti.init(arch=...)``, build model and receiver arrays with
``ti.field(dtype=..., shape=...)`` plus ``from_numpy`` / item assignment,
read results with ``to_numpy()``, and use ``ti.sqrt`` and ``ti.sync`` on the
host.  Here a field is a host NumPy array behind that interface; the solver
classes of this package read model inputs through ``to_numpy()``, and the
time step itself never touches this module.

``install()`` registers the module as ``taichi`` unless a real Taichi is
importable (then nothing changes).
"""
from __future__ import annotations

import importlib.util
import math
import sys

import numpy as np

# arch tokens: accepted and ignored (the step runs on the MI355X regardless)
cpu, gpu, cuda, vulkan, metal, opengl, amdgpu, x64, arm64 = (
    "cpu", "gpu", "cuda", "vulkan", "metal", "opengl", "amdgpu", "x64", "arm64")
f16, f32, f64 = np.float16, np.float32, np.float64
i8, i16, i32, i64 = np.int8, np.int16, np.int32, np.int64
u8, u16, u32, u64 = np.uint8, np.uint16, np.uint32, np.uint64
float32, float64, int32, int64 = f32, f64, i32, i64

_DTYPES = {float: np.float32, int: np.int32}  # Taichi's defaults for Python types


def init(arch=None, **kwargs) -> None:
    return None


def sync() -> None:
    """Wait for the GPU (the solvers' launches are asynchronous)."""
    import torch
    if torch.cuda.is_available():
        torch.cuda.synchronize()


def sqrt(x):
    return np.sqrt(x) if isinstance(x, np.ndarray) else math.sqrt(x)


class Field:
    """``ti.field`` on the host: zero-initialised, fixed dtype and shape."""

    def __init__(self, dtype, shape):
        self.dtype = np.dtype(_DTYPES.get(dtype, dtype))
        self.shape = (shape,) if isinstance(shape, int) else tuple(int(s) for s in shape)
        self._a = np.zeros(self.shape, dtype=self.dtype)

    def from_numpy(self, arr) -> None:
        a = np.asarray(arr)
        if a.shape != self.shape:
            raise ValueError(f"shape {a.shape} does not match field shape {self.shape}")
        self._a[...] = a.astype(self.dtype, copy=False)

    def to_numpy(self) -> np.ndarray:
        return self._a.copy()

    def fill(self, value) -> None:
        self._a.fill(value)

    def __getitem__(self, idx):
        v = self._a[idx]
        return v.item() if np.ndim(v) == 0 else v

    def __setitem__(self, idx, value):
        self._a[idx] = value

    def __array__(self, dtype=None, copy=None):
        return self._a.astype(dtype) if dtype is not None else self._a

    def __repr__(self):
        return f"Field(shape={self.shape}, dtype={self.dtype}) [host stand-in for ti.field]"


def field(dtype, shape, **kwargs) -> Field:
    return Field(dtype, shape)


def install() -> bool:
    """Register this module as ``taichi`` if no real Taichi is importable;
    returns True if it did."""
    if "taichi" in sys.modules or importlib.util.find_spec("taichi") is not None:
        return False
    sys.modules["taichi"] = sys.modules[__name__]
    return True
