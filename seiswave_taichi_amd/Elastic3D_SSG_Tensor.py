"""Drop-in for ``src/Elastic3D_SSG_Tensor.py`` (class ``ElasticWAVE``).

Same constructor (src/Elastic3D_SSG_Tensor.py:9-28) with receiver index fields
``rsx, rsy, rsz`` of shape (1, N) and a (3, 3) moment tensor ``MT``;
``SetADEPML3D`` (:361-559), ``update_SSG(nt)`` (:166-345), ``AEA_init``
(:151-157) and ``AEA`` (:159-163, working here).  Step order: moment-tensor injection
``s_ab[isx,isy,isz] += Ricker2(nt, dt, 0.03, f0, src_scale)*(-M_ab)`` for the
upper triangle of MT (:186-193), velocity sweep, stress sweep, then
``data[nt, r] = vx[rsx[0,r], rsy[0,r], rsz[0,r]]`` (:344-345).
"""
from __future__ import annotations

import numpy as np
import torch

from . import _lib as L
from ._classes3d import ElasticWAVE3DBase, resolve_source
from .coeffs import ricker2
from .field import as_numpy


def mt_components(MT) -> np.ndarray:
    """(Mxx, Myy, Mzz, Mxy, Mxz, Myz) from a (3,3) field, upper triangle as the
    reference reads it (src/Elastic3D_SSG_Tensor.py:187-193)."""
    m = np.asarray(as_numpy(MT), dtype=np.float32).reshape(3, 3)
    return np.array([m[0, 0], m[1, 1], m[2, 2], m[0, 1], m[0, 2], m[1, 2]], dtype=np.float32)


class ElasticWAVE(ElasticWAVE3DBase):
    _CLS = L.SW_CLASS_TENSOR

    def __init__(self, vs, vp, rho, dx: float, dy: float, dz: float, dt: float,
                 isx: int, isy: int, isz: int, rsx, rsy, rsz, MT, src_scale: float, nt: int,
                 accuracy: int, freq=100., fieldtype=None, *, source=None, receivers=None,
                 components=("vx",)):
        rx, ry, rz = (np.asarray(as_numpy(r)).reshape(-1).astype(np.int64) for r in (rsx, rsy, rsz))
        if not (len(rx) == len(ry) == len(rz)):
            raise ValueError("rsx, rsy, rsz must hold the same number of receivers")
        self.rsx, self.rsy, self.rsz, self.MT = rsx, rsy, rsz, MT
        self.src_scale = src_scale
        self.datasize = len(rx)
        mt = mt_components(MT)
        amps = ricker2(np.arange(nt), dt, 0.03, freq, src_scale)
        kind, cell, amps, mt = resolve_source(source, L.SW_SRC_MOMENT, (isx, isy, isz), amps, mt)
        cells = np.stack([rx, ry, rz], axis=1) if receivers is None else receivers
        shape = (nt, len(cells)) if len(components) == 1 else None
        self._setup(vs, vp, rho, dx, dy, dz, dt, isx, isy, isz, nt, accuracy, freq, fieldtype,
                    kind, cell, amps, mt, cells, components, shape)

    def update_SSG(self, nt: int):
        """One reference time step with step index ``nt``."""
        self._run(int(nt), 1)

    def AEA_init(self):
        """src/Elastic3D_SSG_Tensor.py:151-157: lam = 0 and rho halved on planes
        k < accuracy.  The reference mutates the caller's rho field; here the
        device copy of rho is mutated, and so is the caller's array when it is
        a NumPy array or a torch tensor."""
        e = self._e
        g = e.descriptor()
        L.check(e.lib.sw_ssg3d_free_surface_init(g, L.fptr(e.rho_t), L.fptr(e.lam_t), e.stream()),
                "sw_ssg3d_free_surface_init")
        src = self._rho_in
        if isinstance(src, np.ndarray):
            src[:, :, :e.star] *= np.float32(0.5)
        elif isinstance(src, torch.Tensor):
            src[:, :, :e.star] *= 0.5

    def AEA(self):
        """src/Elastic3D_SSG_Tensor.py:159-163: szz[:, :, accuracy] = 0.  The
        reference's is a ti.func that update_SSG never calls (its 3-D free
        surface is dead code); here it is a working per-step call, to be made
        before each update_SSG as the 2-D drivers do (Relay_wave.py:55-58), or
        ``run(..., free_surface=True)`` does it inside the native step loop."""
        e = self._e
        L.check(e.lib.sw_ssg3d_free_surface(e.descriptor(), e.stream()), "sw_ssg3d_free_surface")
