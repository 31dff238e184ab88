"""Scalar wavelet helpers that driver scripts import from the reference's
``src.BaseKernel`` (Relay_wave.py:5 and :57 evaluate ``Ricker(i, dt, freq,
src_scale)`` once per step on the host).

``Ricker`` follows src/BaseKernel.py:4-12 (delay t0 = 1/freq, unlike the
1.5/freq of the solver classes' own source wavelet); ``Ricker2`` follows
src/BaseKernel.py:14-20.  Both evaluate in float32 and return a Python float.
"""
from __future__ import annotations

import numpy as np

from .coeffs import ricker2 as _ricker2

_PI2 = np.float32(np.pi ** 2)


def Ricker(nt: int, dt: float, freq: float, src_scale: float) -> float:
    """-src_scale*2*a*tau*exp(-a*tau**2), tau = nt*dt - 1/freq, a = pi**2 freq**2."""
    f, d, s = np.float32(freq), np.float32(dt), np.float32(src_scale)
    t0 = np.float32(1.0) / f
    a = _PI2 * (f * f)
    tau = np.float32(nt) * d - t0
    e = np.float32(np.exp(np.float64((-a) * (tau * tau))))
    return float((((-s) * np.float32(2)) * a) * tau * e)


def Ricker2(nt: int, dt: float, t0: float, freq: float, src_scale: float) -> float:
    """src_scale*(1 - 2*a*tau**2)*exp(-a*tau**2), tau = nt*dt - t0."""
    return float(_ricker2(np.asarray(nt), dt, t0, freq, src_scale))
