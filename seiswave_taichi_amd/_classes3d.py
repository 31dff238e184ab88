"""Common body of the three 3-D drop-in ``ElasticWAVE`` classes."""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from ._engine3d import WAVEFIELDS, Receivers, Source, SSG3DEngine, _check_fieldtype
from .coeffs import diff_coff as _diff_coff
from .coeffs import ricker, ricker2
from .field import Field, as_numpy

_SRC_KINDS = {"none": L.SW_SRC_NONE, "force_vz": L.SW_SRC_FORCE_VZ, "force_vx": L.SW_SRC_FORCE_VX,
              "moment": L.SW_SRC_MOMENT, "explosive": L.SW_SRC_EXPLOSIVE}

# memory-variable names of the reference (src/Elastic3D_SSG_CPML.py:104-122)
# -> (axis, slab array index) in the device layout
MEMORY_SLOTS = {
    "memory_sxx_dx": (0, 0), "memory_sxy_dx": (0, 1), "memory_sxz_dx": (0, 2),
    "memory_dvx_dx": (0, 3), "memory_dvy_dx": (0, 4), "memory_dvz_dx": (0, 5),
    "memory_sxy_dy": (1, 0), "memory_syy_dy": (1, 1), "memory_syz_dy": (1, 2),
    "memory_dvy_dy": (1, 3), "memory_dvx_dy": (1, 4), "memory_dvz_dy": (1, 5),
    "memory_sxz_dz": (2, 0), "memory_syz_dz": (2, 1), "memory_szz_dz": (2, 2),
    "memory_dvz_dz": (2, 3), "memory_dvx_dz": (2, 4), "memory_dvy_dz": (2, 5),
}

SPLIT_NAMES = ("vx_x", "vx_y", "vx_z", "vy_x", "vy_y", "vy_z", "vz_x", "vz_y", "vz_z",
               "sxx_x", "sxx_y", "sxx_z", "syy_x", "syy_y", "syy_z", "sxy_x", "sxy_y",
               "sxz_x", "sxz_z", "syz_y", "syz_z", "szz_x", "szz_y", "szz_z")


class ElasticWAVE3DBase:
    """State and methods shared by the 3-D classes.

    Subclasses set ``_CLS`` and call ``_setup`` from their reference-signature
    constructor.  Extensions beyond the reference (keyword-only in the
    subclasses): ``source`` overrides the source (``kind`` one of "force_vz",
    "force_vx", "moment", "explosive", "none"; ``cell``; ``mt``; ``wavelet`` as
    an array of per-step amplitudes); ``receivers`` an (N,3) array of cells;
    ``components`` the recorded fields.
    """

    _CLS = L.SW_CLASS_CPML

    def _setup(self, vs, vp, rho, dx, dy, dz, dt, isx, isy, isz, nt, accuracy, freq, fieldtype,
               src_kind, src_cell, amps, mt, rec_cells, components, data_shape):
        _check_fieldtype(fieldtype)
        e = SSG3DEngine(vs, vp, rho, dx, dy, dz, dt, accuracy, nt, self._CLS)
        self._e = e
        self.vs, self.vp = vs, vp
        self._rho_in = rho
        self.gridsize = e.gridsize
        self.dx, self.dy, self.dz, self.dt = dx, dy, dz, dt
        self.xmin = self.ymin = self.zmin = 0
        self.star = e.star
        self.c = e.c
        self.xmax = self.xmin + dx * (self.gridsize[0] - 1 - 2 * accuracy)
        self.ymax = self.zmin + dy * (self.gridsize[1] - 1 - 2 * accuracy)  # Q9: computed from zmin (0)
        self.zmax = self.zmin + dz * (self.gridsize[2] - 1 - 2 * accuracy)
        self.xmin_pml, self.xmax_pml = self.xmin, self.xmax
        self.zmin_pml, self.zmax_pml = self.zmin, self.zmax
        self.f0 = freq
        self.isx, self.isy, self.isz = isx, isy, isz
        self.nt = nt
        self.mu = Field(e.mu_t)
        self.lam = Field(e.lam_t)
        self.rho = rho
        for n in WAVEFIELDS:
            setattr(self, n, e.field(n))
        self._src = Source(e, src_kind, src_cell, amps, mt)
        self._rec = Receivers(e, rec_cells, nt, components)
        d = self._rec.data_t
        self.data = Field(d.view(*data_shape) if data_shape is not None else d, reversed_axes=False)

    # -------------------------------------------------------------- setup
    def SetADEPML3D(self, pml_surface, parameter: dict):
        """src/Elastic3D_SSG_CPML.py:324-522."""
        self._e.set_pml(pml_surface, parameter)
        ax, ay, az = self._e.pml
        self.xmin_pml, self.xmax_pml = ax.min_pml, ax.max_pml
        self.ymin_pml, self.ymax_pml = ay.min_pml, ay.max_pml
        self.zmin_pml, self.zmax_pml = az.min_pml, az.max_pml
        for name, a in zip("xyz", (ax, ay, az)):
            for ref, arr in (("pml_{}", a.d), ("pml_{}_half", a.d_h), ("alpha_{}", a.alpha),
                             ("alpha_{}_half", a.alpha_h), ("k_{}", a.k), ("k_{}_half", a.k_h),
                             ("b_{}", a.b), ("b_{}_half", a.b_h), ("a_{}", a.a), ("a_{}_half", a.a_h)):
                setattr(self, ref.format(name), arr)

    # -------------------------------------------------------------- steps
    def _run(self, it0: int, nsteps: int, flags: int = 0):
        e = self._e
        g = e.descriptor()
        if it0 < 0 or it0 + nsteps > self.nt:
            raise IndexError(f"steps {it0}..{it0 + nsteps - 1} outside 0..{self.nt - 1}")
        L.check(e.lib.sw_ssg3d_run_ex(g, self._src.s, self._rec.r, it0, nsteps, flags, e.stream()),
                "sw_ssg3d_run_ex")

    def run(self, nsteps: int | None = None, it0: int = 0, free_surface: bool = False):
        """Extension: ``nsteps`` reference steps from ``it0`` in one native call.
        ``free_surface``: szz = 0 on plane k = accuracy before every step (the
        3-D counterpart of the reference's 2-D AEA() before each step;
        AEA_init() sets the material side)."""
        self._run(it0, self.nt - it0 if nsteps is None else nsteps,
                  L.SW_RUN_FREE_SURFACE if free_surface else 0)

    def reset(self):
        """Extension: zero the wavefields, boundary state and record."""
        self._e.reset()
        self._rec.data_t.zero_()

    def synchronize(self):
        torch.cuda.synchronize(self._e.device)

    # ------------------------------------------------------ introspection
    def memory_variable(self, name: str) -> np.ndarray:
        """Full-grid (nx,ny,nz) view of a reference CPML memory variable."""
        e = self._e
        if e.mem is None:
            raise AttributeError("this class keeps split fields, not CPML memory variables")
        axis, slot = MEMORY_SLOTS[name]
        full = np.zeros((e.nz, e.ny, e.nx), dtype=np.float32)
        ax = e.pml[axis]
        slab = e.mem[axis][slot].cpu().numpy()
        n = e.gridsize[axis]
        idx = np.r_[0:ax.lo, ax.hi:n]
        if axis == 0:
            full[:, :, idx] = slab
        elif axis == 1:
            full[:, idx, :] = slab
        else:
            full[idx, :, :] = slab
        return np.ascontiguousarray(full.transpose(2, 1, 0))

    def split_field(self, name: str) -> np.ndarray:
        e = self._e
        if e.split is None:
            raise AttributeError("this class keeps CPML memory variables, not split fields")
        comp = e.shell.component(e.split.cpu().numpy(), SPLIT_NAMES.index(name))
        return np.ascontiguousarray(e.shell.unpack(comp).transpose(2, 1, 0))

    @staticmethod
    def diff_coff(order: int):
        return _diff_coff(order)


def resolve_source(source, default_kind, default_cell, default_amps, mt=None):
    """Apply the keyword-only ``source`` override of the 3-D/2-D classes."""
    if source is None:
        return default_kind, default_cell, default_amps, mt
    kind = _SRC_KINDS[source.get("kind", "force_vz")]
    cell = source.get("cell", default_cell)
    amps = np.asarray(source.get("wavelet", default_amps), dtype=np.float32)
    return kind, cell, amps, source.get("mt", mt)


__all__ = ["ElasticWAVE3DBase", "resolve_source", "ricker", "ricker2", "as_numpy", "MEMORY_SLOTS",
           "SPLIT_NAMES"]
