"""Drop-in for ``src/Elastic2D_sparse.py`` (class ``ElasticWAVE``).

The reference class is the 2-D CPML solver of src/Elastic2D_CPML.py with its
eight memory variables in a bitmasked Taichi struct field
(src/Elastic2D_sparse.py:82-94): only the cells its PML branch writes are
activated, and an inactive cell reads 0, so the values are those of dense
memory variables that stay 0 outside the PML.  Here they are the dense device
arrays of the 2-D solver (a 2-D grid's memory variables are a few MB), and the
sweeps are the same HIP kernels.  What the reference class does differently,
and this class reproduces:

* ``lam = rho*(vp**2 - vs**2)`` (src/Elastic2D_sparse.py:105-113: the factor 2
  on ``vs**2`` of src/Elastic2D_CPML.py is missing), in NumPy f32;
* ``update_SSG_VS`` adds ``dt*Ricker(nt, dt, f0, 1)/rho`` at (isx, isz)
  (src/Elastic2D_sparse.py:171), not Ricker2 with t0 = 1.5/f0;
* ``update_SSG_SV`` and ``update_RSG`` read ``self.memory_sxx_dx`` and the
  like (:226, :308), fields the class never creates, so the reference raises
  AttributeError when they are first called; so does this class.

Constructor, ``SetADEPML2D``, ``AEA_init``/``AEA``, ``diff_coff`` and
``diff_coff_op`` are the 2-D CPML class's (src/Elastic2D_sparse.py:11-30,
:116-129, :355-390, :392-...).
"""
from __future__ import annotations

import numpy as np

from . import _lib as L
from .Elastic2D_CPML import ElasticWAVE as _CPML2D
from .coeffs import ricker
from .field import Field, as_numpy, to_device_layout


class ElasticWAVE(_CPML2D):
    def __init__(self, vs, vp, rho, dx: float, dz: float, dt: float, isx: int, isz: int,
                 rsx: int, rsz: int, nt: int, accuracy=3, freq=100., fieldtype=None, *,
                 source=None, receivers=None, components=("vx", "vz")):
        super().__init__(vs, vp, rho, dx, dz, dt, isx, isz, rsx, rsz, nt, accuracy, freq, fieldtype,
                         source=source, receivers=receivers, components=components)
        vs_np, vp_np, rho_np = (np.asarray(as_numpy(a), dtype=np.float32) for a in (vs, vp, rho))
        lam = rho_np * (vp_np ** 2 - vs_np ** 2)
        self.lam_t = to_device_layout(lam.astype(np.float32), self.device)
        self.lam = Field(self.lam_t)
        self._src_vs = self._source(source, L.SW_SRC_FORCE_VZ, ricker(np.arange(nt), dt, freq, 1.0))

    def update_SSG_SV(self, nt: int):
        raise AttributeError("'ElasticWAVE' object has no attribute 'memory_sxx_dx' "
                             "(src/Elastic2D_sparse.py:226: update_SSG_SV refers to memory fields "
                             "the sparse class never creates)")

    def update_RSG(self, nt: int, source: float):
        raise AttributeError("'ElasticWAVE' object has no attribute 'memory_sxx_dx' "
                             "(src/Elastic2D_sparse.py:308: update_RSG refers to memory fields "
                             "the sparse class never creates)")

    def run(self, scheme="VS", nsteps=None, it0=0):
        """Extension: ``nsteps`` VS steps from ``it0`` in one native call."""
        if scheme != "VS":
            raise ValueError("the sparse class steps the VS scheme only")
        super().run("VS", nsteps, it0)
