"""Drop-in for ``src/Elastic2D_CPML.py`` (class ``ElasticWAVE``).

Same constructor (src/Elastic2D_CPML.py:11-25), ``SetADEPML2D`` (:392-528),
``update_SSG_VS(nt)`` (:127-200), ``update_SSG_SV(nt)`` (:202-274),
``update_RSG(nt, source)`` (:276-353), ``AEA_init``/``AEA`` (:112-125),
``diff_coff`` (:355-366) and ``diff_coff_op`` (:368-390); the sweeps run as HIP
kernels.  Step contents:

* VS: velocity, ``vz[isx,isz] += dt*Ricker2(nt, dt, 1.5/f0, f0, 1)/rho``,
  stress, ``data[nt] = vz[rsx, rsz]``;
* SV: velocity, ``vx[isx,isz] += dt*Ricker(nt, dt, f0, 1)/rho``, stress,
  ``data[nt] = vx[rsx, rsz]``;
* RSG: velocity, ``vz[isx+a, isz+b] += dt*source/(8*rho)`` for a, b in
  {-1, 0}, stress; no record.

Extensions (keyword-only): ``source`` (kind "explosive" / "force_vz" /
"force_vx" / "moment", ``cell``, ``mt`` = (Mxx, 0, Mzz, 0, Mxz, 0),
``wavelet`` = per-step amplitudes) and ``receivers`` ((N, 2) cells) with
``components``; then ``data`` has shape (nt, ncomp, N) and every step records.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from ._classes3d import resolve_source
from ._engine3d import _check_fieldtype, gpu_device
from .coeffs import diff_coff as _diff_coff
from .coeffs import diff_coff_op as _diff_coff_op
from .coeffs import ricker, ricker2
from .field import Field, as_numpy, to_device_layout
from .pml import axis_pml, check_parameter

FIELDS2D = ("vx", "vz", "sxx", "szz", "sxz")


class ElasticWAVE:
    def __init__(self, vs, vp, rho, dx: float, dz: float, dt: float, isx: int, isz: int,
                 rsx: int, rsz: int, nt: int, accuracy=3, freq=100., fieldtype=None, *,
                 source=None, receivers=None, components=("vx", "vz")):
        _check_fieldtype(fieldtype)
        self.lib = L.load()
        self.device = dev = gpu_device()
        shape = tuple(int(v) for v in np.shape(vs))
        if len(shape) != 2:
            raise ValueError(f"2-D solver needs (nx, nz) model arrays, got shape {shape}")
        self.vs, self.vp, self._rho_in = vs, vp, rho
        self.gridsize = shape
        self.nx, self.nz = shape
        self.dx, self.dz, self.dt = dx, dz, dt
        self.xmin = self.zmin = 0
        self.star = int(accuracy)
        if not 1 <= self.star <= L.SW_MAX_ORDER:
            raise ValueError(f"accuracy must be in 1..{L.SW_MAX_ORDER}")
        if min(shape) < 2 * self.star + 2:
            raise ValueError(f"grid {shape} too small for accuracy {accuracy}")
        self.c = _diff_coff(self.star)
        self.xmax = self.xmin + dx * (shape[0] - 1 - 2 * accuracy)
        self.zmax = self.zmin + dz * (shape[1] - 1 - 2 * accuracy)
        self.xmin_pml, self.xmax_pml, self.zmin_pml, self.zmax_pml = self.xmin, self.xmax, self.zmin, self.zmax
        vs_np, vp_np, rho_np = (np.asarray(as_numpy(a), dtype=np.float32) for a in (vs, vp, rho))
        mu = rho_np * vs_np ** 2
        lam = rho_np * (vp_np ** 2 - 2 * vs_np ** 2)
        self.rho_t = to_device_layout(rho_np, dev)
        self.mu_t = to_device_layout(mu.astype(np.float32), dev)
        self.lam_t = to_device_layout(lam.astype(np.float32), dev)
        self.mu, self.lam = Field(self.mu_t), Field(self.lam_t)
        self.rho = rho
        self.f0, self.isx, self.isz, self.rsx, self.rsz, self.nt = freq, isx, isz, rsx, rsz, nt
        self.f = {n: torch.zeros((self.nz, self.nx), dtype=torch.float32, device=dev) for n in FIELDS2D}
        for n in FIELDS2D:
            setattr(self, n, Field(self.f[n]))
        self.mem_t = torch.zeros((10, self.nz, self.nx), dtype=torch.float32, device=dev)
        self.pml = None
        steps = np.arange(nt)
        t0 = 1.5 / freq
        self._src_vs = self._source(source, L.SW_SRC_FORCE_VZ, ricker2(steps, dt, t0, freq, 1.0))
        self._src_sv = self._source(source, L.SW_SRC_FORCE_VX, ricker(steps, dt, freq, 1.0))
        self._rsg_amp = torch.zeros(nt, dtype=torch.float32, device=dev)
        self._src_rsg = self._make_source(L.SW_SRC_FORCE_VZ, (isx, isz), self._rsg_amp, None)
        if receivers is None:
            cells = np.array([[rsx, rsz]])
            self.data_t = torch.zeros((nt, 1, 1), dtype=torch.float32, device=dev)
            self._rec_vs = self._receivers(cells, ("vz",), self.data_t)
            self._rec_sv = self._receivers(cells, ("vx",), self.data_t)
            self.data = Field(self.data_t.view(nt), reversed_axes=False)
        else:
            cells = np.asarray(receivers, dtype=np.int64).reshape(-1, 2)
            self.data_t = torch.zeros((nt, len(components), len(cells)), dtype=torch.float32, device=dev)
            self._rec_vs = self._rec_sv = self._receivers(cells, tuple(components), self.data_t)
            self.data = Field(self.data_t, reversed_axes=False)

    # ---------------------------------------------------------------- plumbing
    def _source(self, source, default_kind, amps):
        kind, cell, amps, mt = resolve_source(source, default_kind, (self.isx, self.isz), amps)
        amp_t = torch.from_numpy(np.ascontiguousarray(amps, dtype=np.float32)).to(self.device)
        return self._make_source(kind, cell, amp_t, mt)

    def _make_source(self, kind, cell, amp_t, mt):
        ix, iz = (int(v) for v in cell)
        if not (0 <= ix < self.nx and 0 <= iz < self.nz):
            raise IndexError(f"source cell {cell} outside grid {self.gridsize}")
        s = L.SwSource()
        s.kind, s.ix, s.iy, s.iz = kind, ix, 0, iz
        for i, v in enumerate(np.zeros(6) if mt is None else np.asarray(mt, dtype=np.float32)):
            s.mt[i] = float(v)
        s.amp, s.namp = L.fptr(amp_t), amp_t.numel()
        return (s, amp_t)

    def _receivers(self, cells, comps, data_t):
        cells = np.asarray(cells, dtype=np.int64).reshape(-1, 2)
        if len(cells) and (cells[:, 0].min() < 0 or cells[:, 0].max() >= self.nx or
                           cells[:, 1].min() < 0 or cells[:, 1].max() >= self.nz):
            raise IndexError(f"receiver cells outside grid {self.gridsize}")
        idx_t = torch.from_numpy((cells[:, 1] * self.nx + cells[:, 0]).astype(np.int32)).to(self.device)
        r = L.SwReceivers()
        r.nrec, r.ncomp, r.nt = len(cells), len(comps), data_t.shape[0]
        for i, cn in enumerate(comps):
            r.comp[i] = L.FIELD_CODE[cn]
        r.idx, r.data = L.i32ptr(idx_t), L.fptr(data_t)
        return (r, idx_t)

    def _descriptor(self, scheme, need_pml=True):
        if self.pml is None and need_pml:
            raise RuntimeError("SetADEPML2D must be called before the first time step")
        # built once per scheme and PML setting: the pointers are to tensors that
        # live as long as the instance
        cache = self.__dict__.setdefault("_desc_cache", {})
        if self.pml is not None and scheme in cache:
            return cache[scheme]
        g = L.SwSSG2D()
        g.nx, g.nz, g.order, g.scheme = self.nx, self.nz, self.star, scheme
        g.dx, g.dz, g.dt = float(self.dx), float(self.dz), float(self.dt)
        for i, v in enumerate(self.c):
            g.c[i] = float(v)
        g.rho, g.lam, g.mu = L.fptr(self.rho_t), L.fptr(self.lam_t), L.fptr(self.mu_t)
        for n in FIELDS2D:
            setattr(g, n, L.fptr(self.f[n]))
        for name, (coeff, flags), n in zip(("px", "pz"), self.pml_dev if self.pml else (), (self.nx, self.nz)):
            a = L.SwAxis()
            base = coeff.data_ptr()
            for r, fname in enumerate(("a", "b", "k", "a_h", "b_h", "k_h", "d", "d_h")):
                setattr(a, fname, C.cast(C.c_void_p(base + 4 * r * n), L._fp))
            a.in_ = C.cast(C.c_void_p(flags.data_ptr()), L._u8p)
            a.in_h = C.cast(C.c_void_p(flags.data_ptr() + n), L._u8p)
            a.lo, a.hi = 0, n
            setattr(g, name, a)
        g.mem = L.fptr(self.mem_t)
        if self.pml is not None:
            cache[scheme] = g
        return g

    def _stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    def _step(self, scheme, src, rec, it, nsteps=1):
        if it < 0 or it + nsteps > self.nt:
            raise IndexError(f"steps {it}..{it + nsteps - 1} outside 0..{self.nt - 1}")
        g = self._descriptor(scheme)
        L.check(self.lib.sw_ssg2d_run(g, src[0], None if rec is None else rec[0], it, nsteps, self._stream()),
                "sw_ssg2d_run")

    # ------------------------------------------------------------------ API
    def SetADEPML2D(self, pml_surface, parameter: dict):
        """src/Elastic2D_CPML.py:392-528 (pml_surface: x-left, x-right, z-up, z-bottom)."""
        check_parameter(parameter, ("pml_x_thick", "pml_z_thick"))
        if len(pml_surface) != 4:
            raise ValueError("pml_surface needs 4 flags: x-left, x-right, z-up, z-bottom")
        p = parameter
        ax = axis_pml(self.nx, self.star, self.dx, self.dt, p["pml_x_thick"], p["vp_max"], p["Rcoef"],
                      p["alpha_max_pml"], p["kmax_pml"], p["theta"], pml_surface[0], pml_surface[1])
        az = axis_pml(self.nz, self.star, self.dz, self.dt, p["pml_z_thick"], p["vp_max"], p["Rcoef"],
                      p["alpha_max_pml"], p["kmax_pml"], p["theta"], pml_surface[2], pml_surface[3])
        self.pml = (ax, az)
        self._desc_cache = {}
        self.pml_dev = []
        for a in (ax, az):
            coeff = torch.from_numpy(np.stack([a.a, a.b, a.k, a.a_h, a.b_h, a.k_h, a.d, a.d_h])).to(self.device)
            flags = torch.from_numpy(np.stack([a.inside, a.inside_h])).to(self.device)
            self.pml_dev.append((coeff.contiguous(), flags.contiguous()))
        self.xmin_pml, self.xmax_pml = ax.min_pml, ax.max_pml
        self.zmin_pml, self.zmax_pml = az.min_pml, az.max_pml
        for name, a in (("x", ax), ("z", az)):
            for ref, arr in (("pml_{}", a.d), ("pml_{}_half", a.d_h), ("alpha_{}", a.alpha),
                             ("alpha_{}_half", a.alpha_h), ("k_{}", a.k), ("k_{}_half", a.k_h),
                             ("b_{}", a.b), ("b_{}_half", a.b_h), ("a_{}", a.a), ("a_{}_half", a.a_h)):
                setattr(self, ref.format(name), arr)

    def update_SSG_VS(self, nt: int):
        self._step(L.SW_2D_VS, self._src_vs, self._rec_vs, int(nt))

    def update_SSG_SV(self, nt: int):
        self._step(L.SW_2D_SV, self._src_sv, self._rec_sv, int(nt))

    def update_RSG(self, nt: int, source: float):
        self._rsg_amp[int(nt)] = float(source)
        self._step(L.SW_2D_RSG, self._src_rsg, None, int(nt))

    def run(self, scheme="VS", nsteps=None, it0=0):
        """Extension: ``nsteps`` VS or SV steps from ``it0`` in one native call."""
        n = self.nt - it0 if nsteps is None else nsteps
        if scheme == "VS":
            self._step(L.SW_2D_VS, self._src_vs, self._rec_vs, it0, n)
        elif scheme == "SV":
            self._step(L.SW_2D_SV, self._src_sv, self._rec_sv, it0, n)
        else:
            raise ValueError("run() supports the VS and SV schemes")

    def AEA_init(self):
        """src/Elastic2D_CPML.py:113-119: lam = 0, rho halved on planes i < accuracy
        (the caller's NumPy array or torch tensor is mutated too, as the reference
        mutates the caller's rho field)."""
        g = self._descriptor(L.SW_2D_VS, need_pml=False)
        L.check(self.lib.sw_ssg2d_free_surface_init(g, self._stream()), "sw_ssg2d_free_surface_init")
        src = self._rho_in
        if isinstance(src, np.ndarray):
            src[:self.star, :] *= np.float32(0.5)
        elif isinstance(src, torch.Tensor):
            src[:self.star, :] *= 0.5

    def AEA(self):
        """src/Elastic2D_CPML.py:121-125: sxx[accuracy, :] = 0."""
        g = self._descriptor(L.SW_2D_VS, need_pml=False)
        L.check(self.lib.sw_ssg2d_free_surface(g, self._stream()), "sw_ssg2d_free_surface")

    def reset(self):
        for t in self.f.values():
            t.zero_()
        self.mem_t.zero_()
        self.data_t.zero_()

    def synchronize(self):
        torch.cuda.synchronize(self.device)

    @staticmethod
    def diff_coff(order: int):
        return _diff_coff(order)

    @staticmethod
    def diff_coff_op(order: int):
        return _diff_coff_op(order)
