"""Stencil coefficients and source wavelets (host side).

* ``diff_coff`` restates src/Elastic3D_SSG_CPML.py:311-322 (the same static
  method exists in every ElasticWAVE class): solve ``A c = e0`` with
  ``A[i, j] = (2j+1)**(2i+1)`` in f64 and store the result as f32.
* ``diff_coff_op`` returns the literal "optimised" coefficients of
  src/Elastic2D_CPML.py:368-390.
* ``ricker`` / ``ricker2`` evaluate src/BaseFun.py:4-11 / :13-19 for a whole
  run at once in f32 arithmetic, term by term in the reference's order; the
  exponential is evaluated in f64 and rounded once to f32 (a correctly
  rounded f32 exp in all but double-rounding corner cases).  The kernels read
  one amplitude per step from a device array.
"""
from __future__ import annotations

import numpy as np

# pi**2 is folded by Python in f64 before the kernel sees it (src/BaseFun.py:7).
_PI2 = np.float32(9.869604401089358)


def diff_coff(order: int) -> np.ndarray:
    if not 1 <= int(order) <= 8:
        raise ValueError(f"accuracy must be in 1..8, got {order}")
    order = int(order)
    b = np.zeros(order)
    b[0] = 1
    A = np.zeros((order, order))
    for i in range(order):
        for j in range(order):
            A[i, j] = (2 * j + 1) ** (2 * i + 1)
    return np.linalg.solve(A, b).astype(np.float32)


_OPT = {
    2: [1.129042, -0.0430142],
    3: [2.081695391597685, -0.01139368190892791, -0.2095028691741802],
    4: [1.231666, -1.041182e-1, 2.063707e-2, -3.570998e-3],
    5: [1.236425, -0.10811, 0.02339911, -0.005061550, 0.0007054313],
}


def diff_coff_op(order: int) -> np.ndarray:
    """src/Elastic2D_CPML.py:368-390 (orders 2..5; other orders give zeros, as there)."""
    c = np.zeros(int(order), dtype=np.float32)
    if int(order) in _OPT:
        c[:] = np.asarray(_OPT[int(order)], dtype=np.float32)
    return c


def ricker(steps, dt: float, freq: float, scale: float) -> np.ndarray:
    """-scale*2*a*tau*exp(-a*tau**2), tau = n*dt - 1.5/f, a = pi**2 f**2 (f32)."""
    n = np.asarray(steps)
    f, d, s = np.float32(freq), np.float32(dt), np.float32(scale)
    t0 = np.float32(1.5) / f
    a = _PI2 * (f * f)
    tau = n.astype(np.float32) * d - t0
    e = np.exp(((-a) * (tau * tau)).astype(np.float64)).astype(np.float32)
    return ((((-s) * np.float32(2)) * a) * tau * e).astype(np.float32)


def ricker2(steps, dt: float, t0: float, freq: float, scale: float) -> np.ndarray:
    """scale*(1 - 2*a*tau**2)*exp(-a*tau**2), tau = n*dt - t0 (f32)."""
    n = np.asarray(steps)
    f, d, s, t00 = np.float32(freq), np.float32(dt), np.float32(scale), np.float32(t0)
    a = _PI2 * (f * f)
    tau = n.astype(np.float32) * d - t00
    tt = tau * tau
    e = np.exp(((-a) * tt).astype(np.float64)).astype(np.float32)
    return ((s * (np.float32(1) - (np.float32(2) * a) * tt)) * e).astype(np.float32)
