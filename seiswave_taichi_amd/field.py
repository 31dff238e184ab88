"""``Field``: the object the drop-in classes expose where the reference
exposes a ``ti.field``.

The reference's drivers use ``.shape``, ``.to_numpy()``, ``.from_numpy()``,
``.fill()`` and element access ``f[i, j, k]`` on its fields
(test3d_homegenous_tensor.py:84-86, test3d_homegenous_SPML.py:74).  A
``Field`` presents the reference index order (x, y, z) over a device tensor
stored x fastest, i.e. with the axes reversed: a reference ``(nx, ny, nz)``
field is a ``(nz, ny, nx)`` contiguous tensor on the GPU.
"""
from __future__ import annotations

import numpy as np
import torch


class Field:
    __slots__ = ("data", "_reversed")

    def __init__(self, data: torch.Tensor, reversed_axes: bool = True):
        self.data = data              # device tensor in device layout
        self._reversed = reversed_axes

    @property
    def shape(self):
        s = tuple(self.data.shape)
        return s[::-1] if self._reversed else s

    @property
    def dtype(self):
        return self.data.dtype

    def _view(self) -> torch.Tensor:
        """The tensor indexed in reference order (a strided view, no copy)."""
        if self._reversed:
            return self.data.permute(*reversed(range(self.data.dim())))
        return self.data

    def to_torch(self) -> torch.Tensor:
        return self._view()

    def to_numpy(self) -> np.ndarray:
        return np.ascontiguousarray(self._view().detach().cpu().numpy())

    def from_numpy(self, arr) -> None:
        a = np.asarray(arr)
        if tuple(a.shape) != self.shape:
            raise ValueError(f"shape {a.shape} does not match field shape {self.shape}")
        t = torch.from_numpy(np.ascontiguousarray(a, dtype=np.float32) if self.data.dtype == torch.float32
                             else np.ascontiguousarray(a))
        if self._reversed:
            t = t.permute(*reversed(range(t.dim())))
        self.data.copy_(t.to(self.data.device, self.data.dtype))

    def fill(self, value) -> None:
        self.data.fill_(value)

    def __getitem__(self, idx):
        v = self._view()[idx]
        return v.item() if v.dim() == 0 else v.detach().cpu().numpy()

    def __setitem__(self, idx, value):
        self._view()[idx] = value

    def __array__(self, dtype=None, copy=None):
        a = self.to_numpy()
        return a.astype(dtype) if dtype is not None else a

    def __repr__(self):
        return f"Field(shape={self.shape}, dtype={self.dtype}, device={self.data.device})"


def as_numpy(x) -> np.ndarray:
    """Host copy of a model input in reference index order.

    Accepts a ``Field``, anything with ``.to_numpy()`` (a Taichi field), a torch
    tensor or an array-like.
    """
    if isinstance(x, Field):
        return x.to_numpy()
    if isinstance(x, torch.Tensor):
        return x.detach().cpu().numpy()
    if hasattr(x, "to_numpy"):
        return np.asarray(x.to_numpy())
    return np.asarray(x)


def to_device_layout(a: np.ndarray, device, dtype=torch.float32) -> torch.Tensor:
    """Reference-order host array -> contiguous device tensor with reversed axes."""
    t = torch.from_numpy(np.ascontiguousarray(a))
    t = t.permute(*reversed(range(t.dim()))).contiguous()
    return t.to(device=device, dtype=dtype).contiguous()
