"""Seismogram modelling with an exact discrete adjoint (2-D VS scheme).

``Elastic2DModeling`` runs the 2-D P-SV step of src/Elastic2D_CPML.py:127-200
(update_SSG_VS) for a set of shots and records multi-component seismograms.
``SeismogramFn`` is a ``torch.autograd.Function``: its forward returns the
seismograms for (vp, vs, rho) and its backward runs the hand-written adjoint
kernels of libseiswave (sw_ssg2d_adjoint_*), giving the gradient of any loss
of the seismograms with respect to vp, vs and rho.

Memory: the forward keeps a checkpoint of the full state every
``checkpoint`` steps; the backward recomputes each segment forward from its
checkpoint, keeping of every step the 15 planes the adjoint stages read
(stresses and their memory variables before the velocity stage, velocities
and theirs after it), then runs the adjoint backwards through it, one native
call per segment (sw_ssg2d_adjoint_segment).

Sources: explosive or moment tensor (added to the stresses before the
velocity sweep; no model dependence), or a force on vz or vx (added after the
velocity sweep as dt*amp/rho[p], src/Elastic2D_CPML.py:170,245): the adjoint
then also accumulates, per shot, amp times the adjoint of the forced velocity
at the source cell, which contributes -dt*sum/rho[p]**2 to d/d rho at p.
"""
from __future__ import annotations

import ctypes as C

import numpy as np
import torch

from . import _lib as L
from ._engine3d import gpu_device
from .coeffs import diff_coff, ricker2
from .pml import axis_pml, check_parameter

NSTATE = 15  # vx vz sxx szz sxz + 10 memory variables
NSAVED = 5   # planes per shot and step the adjoint segment keeps (derivative sums)
MAX_BATCH = 64  # SW_MAX_BATCH
_FIELD_ROW = {"vx": 0, "vz": 1, "sxx": 2, "szz": 3, "sxz": 4}


class Elastic2DModeling:
    """Multi-shot 2-D VS modelling on one GPU.

    shape (nx, nz); shots: list of (ix, iz) source cells; receivers: (N, 2)
    cells; components: recorded fields.  ``source`` is "explosive" or a
    moment tensor (Mxx, Mzz, Mxz).  ``wavelet`` defaults to a Ricker of
    ``freq`` delayed by 1.5/freq.  With ``batch`` (the default) all shots (up
    to 64 at a time) advance together in each launch (sw_ssg2d_*_batch);
    otherwise shot by shot.
    """

    def __init__(self, shape, dx, dz, dt, nt, accuracy, pml: dict, shots, receivers,
                 pml_surface=(True, True, True, True), freq=10.0, components=("vx", "vz"),
                 source="explosive", wavelet=None, checkpoint=50, batch=True):
        check_parameter(pml, ("pml_x_thick", "pml_z_thick"))
        self.lib = L.load()
        self.device = gpu_device()
        self.nx, self.nz = (int(v) for v in shape)
        self.dx, self.dz, self.dt, self.nt, self.order = dx, dz, dt, int(nt), int(accuracy)
        self.c = diff_coff(self.order)
        self.checkpoint = int(checkpoint)
        self.batch = bool(batch)
        dev = self.device
        ax = axis_pml(self.nx, self.order, dx, dt, pml["pml_x_thick"], pml["vp_max"], pml["Rcoef"],
                      pml["alpha_max_pml"], pml["kmax_pml"], pml["theta"], pml_surface[0], pml_surface[1])
        az = axis_pml(self.nz, self.order, dz, dt, pml["pml_z_thick"], pml["vp_max"], pml["Rcoef"],
                      pml["alpha_max_pml"], pml["kmax_pml"], pml["theta"], pml_surface[2], pml_surface[3])
        self._axes = []
        for a in (ax, az):
            coeff = torch.from_numpy(np.stack([a.a, a.b, a.k, a.a_h, a.b_h, a.k_h, a.d, a.d_h])).to(dev).contiguous()
            flags = torch.from_numpy(np.stack([a.inside, a.inside_h])).to(dev).contiguous()
            self._axes.append((coeff, flags))
        self.shots = [tuple(int(v) for v in s) for s in shots]
        for s in self.shots:
            if not (1 <= s[0] < self.nx - 1 and 1 <= s[1] < self.nz - 1):
                raise IndexError(f"shot {s} outside the grid")
        rec = np.asarray(receivers, dtype=np.int64).reshape(-1, 2)
        if rec.size and (rec[:, 0].min() < 0 or rec[:, 0].max() >= self.nx or rec[:, 1].min() < 0
                         or rec[:, 1].max() >= self.nz):
            raise IndexError("receiver outside the grid")
        self.rec_idx = torch.from_numpy((rec[:, 1] * self.nx + rec[:, 0]).astype(np.int32)).to(dev)
        self.nrec = len(rec)
        self.components = tuple(components)
        for cn in self.components:
            if cn not in _FIELD_ROW:
                raise ValueError(f"unknown component {cn}")
        if isinstance(source, str):
            kinds = {"explosive": L.SW_SRC_EXPLOSIVE, "force_vz": L.SW_SRC_FORCE_VZ, "force_vx": L.SW_SRC_FORCE_VX}
            if source not in kinds:
                raise ValueError(f"source {source!r}: explosive, force_vz, force_vx or (Mxx, Mzz, Mxz)")
            self.src_kind, mt = kinds[source], np.zeros(6, np.float32)
        else:
            mxx, mzz, mxz = (float(v) for v in source)
            self.src_kind, mt = L.SW_SRC_MOMENT, np.array([mxx, 0, mzz, 0, mxz, 0], np.float32)
        self.mt = mt
        w = ricker2(np.arange(self.nt), dt, 1.5 / freq, freq, 1.0) if wavelet is None else wavelet
        self.amp = torch.from_numpy(np.ascontiguousarray(w, dtype=np.float32)).to(dev)
        # model-dependent device arrays are set by set_model
        self.rho = self.lam = self.mu = None
        self.force = self.src_kind in (L.SW_SRC_FORCE_VZ, L.SW_SRC_FORCE_VX)

    # -------------------------------------------------------------- plumbing
    def set_model(self, vp: torch.Tensor, vs: torch.Tensor, rho: torch.Tensor):
        """vp, vs, rho: (nx, nz) tensors in reference index order."""
        dl = lambda t: t.detach().to(self.device, torch.float32).t().contiguous()  # noqa: E731
        self.vp_d, self.vs_d, self.rho = dl(vp), dl(vs), dl(rho)
        vs2 = self.vs_d * self.vs_d
        self.mu = self.rho * vs2
        self.lam = self.rho * (self.vp_d * self.vp_d - 2 * vs2)

    def new_state(self, n=None):
        shape = (NSTATE, self.nz, self.nx) if n is None else (n, NSTATE, self.nz, self.nx)
        return torch.zeros(shape, dtype=torch.float32, device=self.device)

    def _segment_buffer(self, steps, shots=1):
        """Work of one adjoint segment call: NSAVED planes per shot and step
        (the forward derivative sums the model gradient reads) and 4 scratch
        planes per shot."""
        P = self.nz * self.nx
        return torch.empty((steps * shots * NSAVED + 4 * shots) * P, dtype=torch.float32, device=self.device)

    def desc(self, state: torch.Tensor) -> L.SwSSG2D:
        g = L.SwSSG2D()
        g.nx, g.nz, g.order, g.scheme = self.nx, self.nz, self.order, L.SW_2D_VS
        g.dx, g.dz, g.dt = float(self.dx), float(self.dz), float(self.dt)
        for i, v in enumerate(self.c):
            g.c[i] = float(v)
        g.rho, g.lam, g.mu = L.fptr(self.rho), L.fptr(self.lam), L.fptr(self.mu)
        base = state.data_ptr()
        plane = 4 * self.nx * self.nz
        for name, row in _FIELD_ROW.items():
            setattr(g, name, C.cast(C.c_void_p(base + row * plane), L._fp))
        g.mem = C.cast(C.c_void_p(base + 5 * plane), L._fp)
        for name, (coeff, flags), n in zip(("px", "pz"), self._axes, (self.nx, self.nz)):
            a = L.SwAxis()
            cb = coeff.data_ptr()
            for r, fname in enumerate(("a", "b", "k", "a_h", "b_h", "k_h", "d", "d_h")):
                setattr(a, fname, C.cast(C.c_void_p(cb + 4 * r * n), L._fp))
            a.in_ = C.cast(C.c_void_p(flags.data_ptr()), L._u8p)
            a.in_h = C.cast(C.c_void_p(flags.data_ptr() + n), L._u8p)
            a.lo, a.hi = 0, n
            setattr(g, name, a)
        return g

    def source(self, shot) -> L.SwSource:
        s = L.SwSource()
        s.kind, s.ix, s.iy, s.iz = self.src_kind, shot[0], 0, shot[1]
        for i, v in enumerate(self.mt):
            s.mt[i] = float(v)
        s.amp, s.namp = L.fptr(self.amp), self.nt
        return s

    def receivers(self, data: torch.Tensor) -> L.SwReceivers:
        r = L.SwReceivers()
        r.nrec, r.ncomp, r.nt = self.nrec, len(self.components), self.nt
        for i, cn in enumerate(self.components):
            r.comp[i] = L.FIELD_CODE[cn]
        r.idx, r.data = L.i32ptr(self.rec_idx), L.fptr(data)
        return r

    def stream(self):
        return C.c_void_p(torch.cuda.current_stream(self.device).cuda_stream)

    # --------------------------------------------------------------- forward
    def forward_shot(self, shot, data: torch.Tensor, checkpoints: torch.Tensor | None = None):
        """Run one shot from a zero state; record into data (nt, ncomp, nrec);
        store the state before steps 0, K, 2K, ... in ``checkpoints``."""
        lib, st = self.lib, self.stream()
        state = self.new_state()
        g = self.desc(state)
        src, rec = self.source(shot), self.receivers(data)
        K = self.checkpoint
        for s0 in range(0, self.nt, K):
            if checkpoints is not None:
                checkpoints[s0 // K].copy_(state)
            n = min(K, self.nt - s0)
            L.check(lib.sw_ssg2d_run(g, src, rec, s0, n, st), "sw_ssg2d_run")
        return state

    def _groups(self):
        n = len(self.shots)
        return [list(range(a, min(a + MAX_BATCH, n))) for a in range(0, n, MAX_BATCH)]

    def _cells(self, group):
        arr = (C.c_int32 * (2 * len(group)))()
        for t, s in enumerate(group):
            arr[2 * t], arr[2 * t + 1] = self.shots[s]
        return arr

    def model(self, vp, vs, rho, keep_checkpoints=False):
        """Seismograms (nshot, nt, ncomp, nrec) and, optionally, the checkpoints."""
        if self.batch:
            return self._model_batch(vp, vs, rho, keep_checkpoints)
        self.set_model(vp, vs, rho)
        nck = (self.nt + self.checkpoint - 1) // self.checkpoint
        data = torch.zeros((len(self.shots), self.nt, len(self.components), self.nrec), dtype=torch.float32,
                           device=self.device)
        cks = []
        for s, shot in enumerate(self.shots):
            ck = self.new_state(nck) if keep_checkpoints else None
            self.forward_shot(shot, data[s], ck)
            cks.append(ck)
        return data, cks

    def _model_batch(self, vp, vs, rho, keep_checkpoints):
        self.set_model(vp, vs, rho)
        lib, st = self.lib, self.stream()
        K, nt = self.checkpoint, self.nt
        nck = (nt + K - 1) // K
        ncomp = len(self.components)
        data = torch.zeros((len(self.shots), nt, ncomp, self.nrec), dtype=torch.float32, device=self.device)
        cks = []
        for group in self._groups():
            state = self.new_state(len(group))
            g = self.desc(state[0])
            src = self.source(self.shots[group[0]])
            rec = self.receivers(data[group[0]])
            cells = self._cells(group)
            ck = self.new_state(nck * len(group)).view(nck, len(group), NSTATE, self.nz, self.nx) \
                if keep_checkpoints else None
            for s0 in range(0, nt, K):
                if ck is not None:
                    ck[s0 // K].copy_(state)
                n = min(K, nt - s0)
                L.check(lib.sw_ssg2d_run_batch(g, len(group), NSTATE * self.nx * self.nz, src, cells, rec,
                                               nt * ncomp * self.nrec, s0, n, st), "sw_ssg2d_run_batch")
            cks.append(ck)
        return data, cks

    def _gradient_batch(self, grad_data, checkpoints):
        lib, st = self.lib, self.stream()
        nz, nx, K = self.nz, self.nx, self.checkpoint
        P = nx * nz
        ncomp = len(self.components)
        out = torch.empty((3, nz, nx), dtype=torch.float32, device=self.device)
        total = None
        g_src = torch.zeros(len(self.shots), dtype=torch.float32, device=self.device)
        for group, ck in zip(self._groups(), checkpoints):
            nb = len(group)
            # per shot: 15 adjoint state planes, 4 scratch, 5 accumulators
            adjb = torch.zeros((nb, 24, nz, nx), dtype=torch.float32, device=self.device)
            q = self._adj_desc(adjb[0, :NSTATE], adjb[0, NSTATE:NSTATE + 4], adjb[0, NSTATE + 4:],
                               g_src[group[0]:] if self.force else None)
            work = self.new_state(nb)
            seg = self._segment_buffer(K, nb)
            g = self.desc(work[0])
            src = self.source(self.shots[group[0]])
            res = grad_data[group[0]:group[-1] + 1].detach().to(torch.float32).contiguous()
            rres = self.receivers(res[0])
            cells = self._cells(group)
            for s0 in reversed(range(0, self.nt, K)):
                n = min(K, self.nt - s0)
                work.copy_(ck[s0 // K])
                L.check(lib.sw_ssg2d_adjoint_segment_batch(g, nb, NSTATE * P, src, cells, q, 24 * P, rres,
                                                           self.nt * ncomp * self.nrec, L.fptr(seg), s0, n, st),
                        "sw_ssg2d_adjoint_segment_batch")
            part = torch.empty((3, nz, nx), dtype=torch.float32, device=self.device)
            L.check(lib.sw_ssg2d_model_gradient_batch(g, q, nb, 24 * P, L.fptr(self.vp_d), L.fptr(self.vs_d),
                                                      L.fptr(part[0]), L.fptr(part[1]), L.fptr(part[2]), st),
                    "sw_ssg2d_model_gradient_batch")
            total = part if total is None else total + part
        out.copy_(total)
        self._force_gradient(out[2], g_src)
        return tuple(out[i].t() for i in range(3))

    def _force_gradient(self, grho, g_src):
        """d/d rho[p] of the forces dt*amp/rho[p]: -dt*g_src/rho[p]**2 per shot
        (grho in device layout [nz][nx])."""
        if not self.force:
            return
        for s, (ix, iz) in enumerate(self.shots):
            grho[iz, ix] += -float(self.dt) * g_src[s] / (self.rho[iz, ix] * self.rho[iz, ix])

    # -------------------------------------------------------------- backward
    def gradient(self, grad_data: torch.Tensor, checkpoints):
        """d loss / d (vp, vs, rho) for d loss / d data = grad_data, as (nx, nz)
        tensors in reference order.  Requires the model set by ``model``."""
        if self.batch:
            return self._gradient_batch(grad_data, checkpoints)
        lib, st = self.lib, self.stream()
        nz, nx, K = self.nz, self.nx, self.checkpoint
        acc = torch.zeros((5, nz, nx), dtype=torch.float32, device=self.device)  # g_b1 g_b4 g_lamh g_muh g_muz
        adj = self.new_state()
        tmp = torch.zeros((4, nz, nx), dtype=torch.float32, device=self.device)
        seg = self._segment_buffer(K)
        work = self.new_state()
        g_src = torch.zeros(len(self.shots), dtype=torch.float32, device=self.device)
        gw = self.desc(work)
        for s, shot in enumerate(self.shots):
            q = self._adj_desc(adj, tmp, acc, g_src[s:] if self.force else None)
            res = grad_data[s].detach().to(torch.float32).contiguous()
            rres = self.receivers(res)
            src = self.source(shot)
            adj.zero_()
            for s0 in reversed(range(0, self.nt, K)):
                n = min(K, self.nt - s0)
                work.copy_(checkpoints[s][s0 // K])
                L.check(lib.sw_ssg2d_adjoint_segment(gw, src, q, rres, L.fptr(seg), s0, n, st),
                        "sw_ssg2d_adjoint_segment")
        out = torch.empty((3, nz, nx), dtype=torch.float32, device=self.device)
        L.check(lib.sw_ssg2d_model_gradient(self.desc(work), q, L.fptr(self.vp_d), L.fptr(self.vs_d),
                                            L.fptr(out[0]), L.fptr(out[1]), L.fptr(out[2]), st),
                "sw_ssg2d_model_gradient")
        self._force_gradient(out[2], g_src)
        return tuple(out[i].t() for i in range(3))

    def _adj_desc(self, adj, tmp, acc, g_src=None) -> "L.SwAdj2D":
        q = L.SwAdj2D()
        if g_src is not None:
            q.g_src = L.fptr(g_src)
        plane = 4 * self.nx * self.nz
        base = adj.data_ptr()
        for name, row in _FIELD_ROW.items():
            setattr(q, name, C.cast(C.c_void_p(base + row * plane), L._fp))
        q.mem = C.cast(C.c_void_p(base + 5 * plane), L._fp)
        q.tmp = L.fptr(tmp)
        for r, name in enumerate(("g_b1", "g_b4", "g_lamh", "g_muh", "g_muz")):
            setattr(q, name, C.cast(C.c_void_p(acc.data_ptr() + r * plane), L._fp))
        return q


class SeismogramFn(torch.autograd.Function):
    """data = F(vp, vs, rho); backward through the hand-written adjoint."""

    @staticmethod
    def forward(ctx, vp, vs, rho, modeling: Elastic2DModeling):
        data, cks = modeling.model(vp, vs, rho, keep_checkpoints=True)
        ctx.modeling = modeling
        ctx.cks = cks
        ctx.model = (vp.detach(), vs.detach(), rho.detach())
        return data

    @staticmethod
    def backward(ctx, grad_data):
        m = ctx.modeling
        m.set_model(*ctx.model)
        gvp, gvs, grho = m.gradient(grad_data, ctx.cks)
        return gvp, gvs, grho, None


def seismograms(vp, vs, rho, modeling: Elastic2DModeling):
    return SeismogramFn.apply(vp, vs, rho, modeling)
