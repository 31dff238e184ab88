"""Wiggle plot of a gather, the helper driver scripts import from the
reference's ``src.wigb`` (test3d_homegenous_tensor.py:5, :95).

``wigb(a, scal, x, z, amx)`` follows src/wigb.py:4-20 in its arguments: ``a``
is [nt, ntraces] (traces are columns), ``scal`` multiplies the traces, ``x``
and ``z`` annotate the trace and time axes (default 1..n), ``amx`` is the
amplitude drawn one trace spacing wide (default: the mean of the traces'
peak amplitudes).  Positive lobes are filled black, time increases downwards.
Returns the matplotlib figure after ``plt.show()``.
"""
from __future__ import annotations

import numpy as np


def wigb(a=None, scal=0, x=0, z=0, amx=0):
    import matplotlib.pyplot as plt

    a = np.random.random((10, 10)) if a is None or (isinstance(a, list) and not a) else np.asarray(a, float)
    nz, nx = a.shape
    if nx <= 1:
        print("ERR:PlotWig: nx has to be more than 1")
        return None
    peaks = np.abs(a).max(axis=0)
    amx = float(peaks.mean()) if not amx else float(amx)
    x = np.arange(1, nx + 1, dtype=float) if np.isscalar(x) and x == 0 else np.asarray(x, float)
    z = np.arange(1, nz + 1, dtype=float) if np.isscalar(z) and z == 0 else np.asarray(z, float)
    scal = 1.0 if not scal else float(scal)
    dx = float(np.median(np.abs(np.diff(x))))
    dz = float(z[1] - z[0])
    print(" PlotWig: data range [%f, %f], plotted max %f \n" % (a.min(), a.max(), amx))
    w = a * (dx / amx if amx else 0.0) * scal
    fig, ax = plt.subplots()
    for i in range(nx):
        if peaks[i] == 0:
            ax.plot([x[i], x[i]], [z[0], z[-1]], color="k", linewidth=1.0)
            continue
        tr = x[i] + w[:, i]
        ax.fill_betweenx(z, x[i], tr, where=w[:, i] > 0, color="k", interpolate=True)
        ax.plot(tr, z, color="k", linewidth=1.0)
    ax.set_xlim(x.min() - 2.0 * dx, x.max() + 2.0 * dx)
    ax.set_ylim(z.min() - dz, z.max() + dz)
    ax.invert_yaxis()
    plt.show()
    return fig
