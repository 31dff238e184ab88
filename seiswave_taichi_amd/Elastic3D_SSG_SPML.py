"""Drop-in for ``src/Elastic3D_SSG_SPML.py`` (class ``ElasticWAVE``).

Same constructor (src/Elastic3D_SSG_SPML.py:8-27, with ``src_scale``),
``SetADEPML3D`` (:408-606) and ``update_SSG_SPML(nt)`` (:160-393): split-field
PML with a Crank-Nicolson step on the damping d inside the PML region, point
lambda and mu for the normal stresses, force
``vz[isx,isy,isz] += dt*Ricker(nt, dt, f0, src_scale)/rho`` (:294) and
``data[nt] = vx[rsx, rsy, rsz]`` (:393).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from ._classes3d import ElasticWAVE3DBase, resolve_source
from .coeffs import ricker


class ElasticWAVE(ElasticWAVE3DBase):
    _CLS = L.SW_CLASS_SPML

    def __init__(self, vs, vp, rho, dx: float, dy: float, dz: float, dt: float,
                 isx: int, isy: int, isz: int, rsx: int, rsy: int, rsz: int, nt: int,
                 accuracy: int, freq=100., src_scale=1., fieldtype=None, *, source=None,
                 receivers=None, components=("vx",)):
        self.src_scale = src_scale
        self.dl = float(np.sqrt(dx ** 2 + dy ** 2 + dz ** 2))
        amps = ricker(np.arange(nt), dt, freq, src_scale)
        kind, cell, amps, mt = resolve_source(source, L.SW_SRC_FORCE_VZ, (isx, isy, isz), amps)
        self.rsx, self.rsy, self.rsz = rsx, rsy, rsz
        cells = [(rsx, rsy, rsz)] if receivers is None else receivers
        shape = (nt,) if receivers is None and len(components) == 1 else None
        self._setup(vs, vp, rho, dx, dy, dz, dt, isx, isy, isz, nt, accuracy, freq, fieldtype,
                    kind, cell, amps, mt, cells, components, shape)

    def update_SSG_SPML(self, nt: int):
        """One reference time step with step index ``nt``."""
        self._run(int(nt), 1)
