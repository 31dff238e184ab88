"""Shared 3-D engine behind the drop-in ElasticWAVE classes.

Holds the device state (torch-ROCm tensors used only as device memory),
builds the ``sw_ssg3d`` descriptor of include/seiswave.h and calls
libseiswave.so through ctypes on torch's current stream.

Reference anchors: constructor state src/Elastic3D_SSG_CPML.py:9-122,
moduli :124-141, SetADEPML3D :324-522, update_SSG :144-309.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from .coeffs import diff_coff
from .field import Field, as_numpy, to_device_layout
from .pml import axis_pml, check_parameter

WAVEFIELDS = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")


def split_free_range(ax):
    """[lo, hi): the indices of an axis where neither the integer nor the half
    point is in the split-field PML (both flags 0); (0, 0) if there are none."""
    free = ~(np.asarray(ax.inside, bool) | np.asarray(ax.inside_h, bool))
    idx = np.nonzero(free)[0]
    if idx.size == 0:
        return 0, 0
    lo, hi = int(idx[0]), int(idx[-1]) + 1
    if not free[lo:hi].all():
        raise ValueError("split-field PML flags are not a pair of slabs")
    return lo, hi


class SplitShell:
    """Shell storage of the split-field PML (include/seiswave.h, sw_ssg3d.split):
    each of the 24 split components keeps only the cells outside the
    split-free box [lo, hi) of every axis, where the reference's full-grid
    split arrays are exactly 0.  Parts: the planes outside [zlo, zhi)
    ([wz][ny][nx]); in the other planes the rows outside [ylo, yhi)
    ([mz][wy][nx]); in those rows the columns outside [xlo, xhi)
    ([mz][my][wx])."""

    def __init__(self, shape_xyz, ranges_xyz):
        nx, ny, nz = (int(v) for v in shape_xyz)
        (self.xlo, self.xhi), (self.ylo, self.yhi), (self.zlo, self.zhi) = ranges_xyz
        self.shape = (nz, ny, nx)
        self.wx = nx - (self.xhi - self.xlo)
        self.wy = ny - (self.yhi - self.ylo)
        self.wz = nz - (self.zhi - self.zlo)
        self.my = self.yhi - self.ylo
        self.mz = self.zhi - self.zlo
        self.OZ = self.wz * ny * nx
        self.OY = self.OZ + self.mz * self.wy * nx
        self.cells = self.OY + self.mz * self.my * self.wx
        self.floats = 1536 * ((self.cells + 63) // 64)

    # component c -> (group offset in a 64-slot block, group width, first component)
    GROUPS = [(0, 4, 0)] * 4 + [(256, 4, 4)] * 4 + [(512, 1, 8)] + [(576, 4, 9)] * 4 + [(832, 4, 13)] * 4 \
        + [(1088, 4, 17)] * 4 + [(1344, 2, 21)] * 2 + [(1472, 1, 23)]

    def index(self, c, slots):
        """Float index in the storage of component c at the given slots."""
        off, w, c0 = self.GROUPS[c]
        slots = np.asarray(slots, np.int64)
        return (slots >> 6) * 1536 + off + w * (slots & 63) + (c - c0)

    def component(self, storage, c) -> np.ndarray:
        """Component c's values at slots 0..S-1 of the storage."""
        storage = np.asarray(storage).reshape(-1)
        if storage.size != self.floats:
            raise ValueError(f"{storage.size} floats, the shell storage holds {self.floats}")
        return storage[self.index(c, np.arange(self.cells))]

    def _slab(self, n, lo, hi):
        return np.r_[0:lo, hi:n]

    def unpack(self, comp) -> np.ndarray:
        """Full [nz][ny][nx] array of one component's values at slots
        0..S-1 (0 inside the split-free box)."""
        nz, ny, nx = self.shape
        comp = np.asarray(comp).reshape(-1)
        if comp.size != self.cells:
            raise ValueError(f"{comp.size} values, the shell holds {self.cells}")
        full = np.zeros(self.shape, dtype=comp.dtype)
        kz = self._slab(nz, self.zlo, self.zhi)
        jy = self._slab(ny, self.ylo, self.yhi)
        ix = self._slab(nx, self.xlo, self.xhi)
        full[kz] = comp[:self.OZ].reshape(self.wz, ny, nx)
        mid = full[self.zlo:self.zhi]
        mid[:, jy, :] = comp[self.OZ:self.OY].reshape(self.mz, self.wy, nx)
        inner = mid[:, self.ylo:self.yhi]
        inner[:, :, ix] = comp[self.OY:].reshape(self.mz, self.my, self.wx)
        return full

    def slot(self, i, j, k) -> int:
        """Index of cell (i, j, k) in a component's shell storage (ValueError
        inside the split-free box)."""
        nz, ny, nx = self.shape
        sl = lambda v, lo, hi: v if v < lo else v - (hi - lo)
        if not (self.zlo <= k < self.zhi):
            return (sl(k, self.zlo, self.zhi) * ny + j) * nx + i
        if not (self.ylo <= j < self.yhi):
            return self.OZ + ((k - self.zlo) * self.wy + sl(j, self.ylo, self.yhi)) * nx + i
        if not (self.xlo <= i < self.xhi):
            return self.OY + ((k - self.zlo) * self.my + (j - self.ylo)) * self.wx + sl(i, self.xlo, self.xhi)
        raise ValueError(f"cell ({i}, {j}, {k}) is in the split-free box")


def _check_fieldtype(fieldtype):
    name = getattr(fieldtype, "__name__", None) or str(fieldtype)
    if fieldtype in (None, np.float32, torch.float32) or "f32" in name or "float32" in name:
        return
    raise NotImplementedError(
        f"fieldtype {fieldtype!r}: the MI355X kernels compute in f32 only (f64 exists in the CPU oracle)")


def gpu_device() -> torch.device:
    if not torch.cuda.is_available():
        raise RuntimeError("seiswave_taichi_amd needs a ROCm GPU: torch.cuda.is_available() is False")
    return torch.device("cuda", torch.cuda.current_device())


def moduli(vs: np.ndarray, vp: np.ndarray, rho: np.ndarray):
    """mu = rho*vs**2, lam = rho*(vp**2 - 2*vs**2) in NumPy f32, as the
    reference computes them (src/Elastic3D_SSG_CPML.py:124-141)."""
    vs, vp, rho = (np.asarray(a, dtype=np.float32) for a in (vs, vp, rho))
    mu = rho * vs ** 2
    lam = rho * (vp ** 2 - 2 * vs ** 2)
    return mu.astype(np.float32), lam.astype(np.float32)


class SSG3DEngine:
    """Device state and launches for one 3-D ElasticWAVE instance."""

    def __init__(self, vs, vp, rho, dx, dy, dz, dt, accuracy, nt, cls):
        _check_dims = np.shape(vs)
        if len(_check_dims) != 3:
            raise ValueError(f"3-D solver needs (nx, ny, nz) model arrays, got shape {_check_dims}")
        self.lib = L.load()
        self.device = gpu_device()
        self.cls = cls
        self.gridsize = tuple(int(v) for v in _check_dims)
        self.nx, self.ny, self.nz = self.gridsize
        self.star = int(accuracy)
        if not 1 <= self.star <= L.SW_MAX_ORDER:
            raise ValueError(f"accuracy must be in 1..{L.SW_MAX_ORDER}, got {accuracy}")
        need = 2 * self.star + 2
        if min(self.gridsize) < need:
            raise ValueError(f"grid {self.gridsize} too small for accuracy {accuracy} (each axis >= {need})")
        self.dx, self.dy, self.dz, self.dt = dx, dy, dz, dt
        self.nt = int(nt)
        self.c = diff_coff(self.star)
        vs_np, vp_np, rho_np = (as_numpy(a) for a in (vs, vp, rho))
        for name, a in (("vs", vs_np), ("vp", vp_np), ("rho", rho_np)):
            if a.shape != self.gridsize:
                raise ValueError(f"{name} has shape {a.shape}, expected {self.gridsize}")
        mu_np, lam_np = moduli(vs_np, vp_np, rho_np)
        dev = self.device
        self.rho_t = to_device_layout(rho_np.astype(np.float32), dev)
        self.mu_t = to_device_layout(mu_np, dev)
        self.lam_t = to_device_layout(lam_np, dev)
        shape = (self.nz, self.ny, self.nx)
        self.f = {n: torch.zeros(shape, dtype=torch.float32, device=dev) for n in WAVEFIELDS}
        self.pml = None
        self._desc = None
        self._keep = []

    # ------------------------------------------------------------- PML
    def set_pml(self, pml_surface, parameter: dict):
        """SetADEPML3D (src/Elastic3D_SSG_CPML.py:324-522)."""
        check_parameter(parameter, ("pml_x_thick", "pml_y_thick", "pml_z_thick"))
        if len(pml_surface) != 6:
            raise ValueError("pml_surface needs 6 flags: x-left, x-right, y-left, y-right, z-up, z-bottom")
        p = parameter
        axes = []
        for n, d, thick, lo, hi in zip(self.gridsize, (self.dx, self.dy, self.dz),
                                       (p["pml_x_thick"], p["pml_y_thick"], p["pml_z_thick"]),
                                       pml_surface[0::2], pml_surface[1::2]):
            axes.append(axis_pml(n, self.star, d, self.dt, thick, p["vp_max"], p["Rcoef"],
                                 p["alpha_max_pml"], p["kmax_pml"], p["theta"], bool(lo), bool(hi)))
        self.pml = axes
        self._desc = None
        dev = self.device
        self.pml_dev = []
        for ax in axes:
            coeff = torch.from_numpy(np.stack([ax.a, ax.b, ax.k, ax.a_h, ax.b_h, ax.k_h, ax.d, ax.d_h])).to(dev)
            flags = torch.from_numpy(np.stack([ax.inside, ax.inside_h])).to(dev)
            self.pml_dev.append((coeff.contiguous(), flags.contiguous()))
        self._alloc_boundary_state()

    def _alloc_boundary_state(self):
        dev = self.device
        nx, ny, nz = self.nx, self.ny, self.nz
        ax, ay, az = self.pml
        if self.cls == L.SW_CLASS_SPML:
            self.shell = SplitShell((nx, ny, nz), [split_free_range(a) for a in (ax, ay, az)])
            self.split = torch.zeros(self.shell.floats, dtype=torch.float32, device=dev)
            self.mem = None
        else:
            self.split = None
            self.mem = (
                torch.zeros((6, nz, ny, ax.width), dtype=torch.float32, device=dev),
                torch.zeros((6, nz, ay.width, nx), dtype=torch.float32, device=dev),
                torch.zeros((6, az.width, ny, nx), dtype=torch.float32, device=dev),
            )

    def reset(self):
        """Zero the wavefields and boundary state (a fresh run, same model)."""
        for t in self.f.values():
            t.zero_()
        if self.split is not None:
            self.split.zero_()
        if self.mem is not None:
            for t in self.mem:
                t.zero_()

    # ------------------------------------------------------- descriptor
    def descriptor(self) -> L.SwSSG3D:
        """The sw_ssg3d of this state (built once per PML setting: every
        pointer in it is to a tensor that lives as long as the engine)."""
        if self.pml is None:
            raise RuntimeError("SetADEPML3D must be called before the first time step "
                               "(the reference raises AttributeError on ymin_pml otherwise)")
        if self._desc is not None:
            return self._desc
        g = L.SwSSG3D()
        g.nx, g.ny, g.nz, g.order, g.cls = self.nx, self.ny, self.nz, self.star, self.cls
        g.dx, g.dy, g.dz, g.dt = float(self.dx), float(self.dy), float(self.dz), float(self.dt)
        for i, v in enumerate(self.c):
            g.c[i] = float(v)
        g.rho, g.lam, g.mu = (L.fptr(t) for t in (self.rho_t, self.lam_t, self.mu_t))
        for n in WAVEFIELDS:
            setattr(g, n, L.fptr(self.f[n]))
        for name, ax, (coeff, flags) in zip(("px", "py", "pz"), self.pml, self.pml_dev):
            a = L.SwAxis()
            base = coeff.data_ptr()
            n = coeff.shape[1]
            for r, fname in enumerate(("a", "b", "k", "a_h", "b_h", "k_h", "d", "d_h")):
                setattr(a, fname, C.cast(C.c_void_p(base + 4 * r * n), L._fp))
            a.in_ = C.cast(C.c_void_p(flags.data_ptr()), L._u8p)
            a.in_h = C.cast(C.c_void_p(flags.data_ptr() + n), L._u8p)
            if self.cls == L.SW_CLASS_SPML:
                a.lo, a.hi = split_free_range(ax)  # the shell's box (SplitShell)
            else:
                a.lo, a.hi = ax.lo, ax.hi
            setattr(g, name, a)
        if self.mem is not None:
            g.mem_x, g.mem_y, g.mem_z = (L.fptr(t) for t in self.mem)
        if self.split is not None:
            g.split = L.fptr(self.split)
        self._desc = g
        return g

    def stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def device_index(self, i, j, k) -> int:
        return (int(k) * self.ny + int(j)) * self.nx + int(i)

    def field(self, name) -> Field:
        return Field(self.f[name])


class Source:
    """Device-side source description (``sw_source``)."""

    def __init__(self, engine, kind, cell, amps: np.ndarray, mt=None):
        self.kind = kind
        self.cell = tuple(int(v) for v in cell)
        self.amp_t = torch.from_numpy(np.ascontiguousarray(amps, dtype=np.float32)).to(engine.device)
        self.mt = np.zeros(6, dtype=np.float32) if mt is None else np.asarray(mt, dtype=np.float32)
        self.s = L.SwSource()
        self.s.kind = kind
        self.s.ix, self.s.iy, self.s.iz = self.cell
        for i, v in enumerate(self.mt):
            self.s.mt[i] = float(v)
        self.s.amp = L.fptr(self.amp_t)
        self.s.namp = len(amps)
        shape = engine.gridsize
        if kind != L.SW_SRC_NONE and not all(0 <= c < n for c, n in zip(self.cell, shape)):
            raise IndexError(f"source cell {self.cell} outside grid {shape}")


class Receivers:
    """Device-side receiver gather (``sw_receivers``): data is [nt][ncomp][nrec]."""

    def __init__(self, engine, cells: np.ndarray, nt: int, components=("vx",)):
        cells = np.asarray(cells, dtype=np.int64).reshape(-1, 3)
        shape = engine.gridsize
        for d in range(3):
            if len(cells) and (cells[:, d].min() < 0 or cells[:, d].max() >= shape[d]):
                raise IndexError(f"receiver cells outside grid {shape}")
        idx = (cells[:, 2] * engine.ny + cells[:, 1]) * engine.nx + cells[:, 0]
        self.cells = cells
        self.components = tuple(components)
        self.idx_t = torch.from_numpy(idx.astype(np.int32)).to(engine.device)
        self.data_t = torch.zeros((nt, len(self.components), len(cells)), dtype=torch.float32,
                                  device=engine.device)
        r = L.SwReceivers()
        r.nrec, r.ncomp, r.nt = len(cells), len(self.components), nt
        for i, cname in enumerate(self.components):
            r.comp[i] = L.FIELD_CODE[cname]
        r.idx = L.i32ptr(self.idx_t)
        r.data = L.fptr(self.data_t)
        self.r = r
