"""Drop-in for ``src/Elastic3D_SSG_CPML.py`` (class ``ElasticWAVE``).

Same constructor (src/Elastic3D_SSG_CPML.py:9-26), ``SetADEPML3D`` (:324-522),
``update_SSG(nt)`` (:144-309) and ``diff_coff`` (:311-322); the time step runs
as HIP kernels on the GPU.  Step order per ``update_SSG``: velocity sweep,
``vz[isx+5, isy+5, isz+5] += dt*Ricker(nt, dt, f0, 1000)/rho`` (:245), stress
sweep, ``data[nt] = vx[rsx, rsy, rsz]`` (:309).  The six Dirichlet loops
(:213-243) write zeros to cells no loop updates and are not run.
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from ._classes3d import ElasticWAVE3DBase, resolve_source
from .coeffs import ricker


class ElasticWAVE(ElasticWAVE3DBase):
    _CLS = L.SW_CLASS_CPML

    def __init__(self, vs, vp, rho, dx: float, dy: float, dz: float, dt: float,
                 isx: int, isy: int, isz: int, rsx: int, rsy: int, rsz: int, nt: int,
                 accuracy: int, freq=100, fieldtype=None, *, source=None, receivers=None,
                 components=("vx",)):
        amps = ricker(np.arange(nt), dt, freq, 1000.0)
        kind, cell, amps, mt = resolve_source(source, L.SW_SRC_FORCE_VZ, (isx + 5, isy + 5, isz + 5), amps)
        self.rsx, self.rsy, self.rsz = rsx, rsy, rsz
        cells = [(rsx, rsy, rsz)] if receivers is None else receivers
        shape = (nt,) if receivers is None and len(components) == 1 else None
        self._setup(vs, vp, rho, dx, dy, dz, dt, isx, isy, isz, nt, accuracy, freq, fieldtype,
                    kind, cell, amps, mt, cells, components, shape)

    def update_SSG(self, nt: int):
        """One reference time step with step index ``nt``."""
        self._run(int(nt), 1)
