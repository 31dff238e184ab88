"""Seeded synthetic earth models.

``layered_folded_model`` builds the heterogeneous full-scale model: eight
folded interfaces ``z_m(x, y) = zbar_m + A_m sin(2 pi x / L_m + phi_m)
sin(2 pi y / M_m + psi_m)``, layer P velocity rising linearly from 1500 m/s
(top, water) to 5500 m/s, a Gaussian-correlated random perturbation of 5 %
standard deviation and the requested correlation length below the water,
``vs = vp / 1.73`` (0 in the water layer) and Gardner's density
``rho = 1741 (vp/1000)^0.25`` as the reference's overthrust driver writes it
(test3d_homegenous_overhust.py:15).  Arrays are (nx, ny, nz) float32 in the
reference index order, z being depth.
"""
from __future__ import annotations

import numpy as np


def gaussian_noise_field(shape, corr_cells: float, rng, coarsen: int = 2) -> np.ndarray:
    """Zero-mean, unit-std Gaussian random field with a Gaussian correlation
    kernel of standard deviation ``corr_cells`` (in cells).  The field is drawn
    on a grid ``coarsen`` times coarser and interpolated trilinearly, which
    keeps the generator fast at 10^8 cells."""
    from scipy import ndimage

    coarsen = max(1, int(coarsen))
    cshape = tuple(max(4, -(-n // coarsen)) for n in shape)
    white = rng.standard_normal(cshape).astype(np.float32)
    f = ndimage.gaussian_filter(white, sigma=corr_cells / coarsen, mode="wrap")
    if coarsen > 1:
        f = ndimage.zoom(f, [n / c for n, c in zip(shape, cshape)], order=1, mode="nearest", grid_mode=True)
        f = f[tuple(slice(0, n) for n in shape)]
    f -= f.mean()
    f /= f.std()
    return f.astype(np.float32)


def layered_folded_model(shape, d: float, seed: int = 0, n_interfaces: int = 8, vp_top: float = 1500.0,
                         vp_bottom: float = 5500.0, perturbation: float = 0.05, corr_len: float = 60.0):
    rng = np.random.default_rng(seed)
    nx, ny, nz = shape
    x = np.arange(nx, dtype=np.float32) * d
    y = np.arange(ny, dtype=np.float32) * d
    z = np.arange(nz, dtype=np.float32) * d
    depth = nz * d
    zbar = np.linspace(0.06, 0.92, n_interfaces) * depth
    spacing = depth / (n_interfaces + 1)
    layer = np.zeros(shape, dtype=np.int8)
    for m in range(n_interfaces):
        amp = (0.15 + 0.25 * rng.random()) * spacing * (0.2 if m == 0 else 1.0)
        lx, ly = rng.uniform(0.3, 1.0, 2) * np.array([nx * d, ny * d])
        px, py = rng.uniform(0, 2 * np.pi, 2)
        surf = zbar[m] + amp * np.outer(np.sin(2 * np.pi * x / lx + px), np.sin(2 * np.pi * y / ly + py))
        layer += (z[None, None, :] >= surf[:, :, None]).astype(np.int8)
    vlayers = np.linspace(vp_top, vp_bottom, n_interfaces + 1).astype(np.float32)
    vp = vlayers[layer]
    water = layer == 0
    pert = gaussian_noise_field(shape, corr_len / d, rng)
    vp = np.where(water, vp, vp * (1.0 + perturbation * pert)).astype(np.float32)
    vs = np.where(water, 0.0, vp / 1.73).astype(np.float32)
    rho = (1741.0 * (vp / 1000.0) ** 0.25).astype(np.float32)
    return vs, vp, rho


def cfl_dt(vp_max: float, d, courant: float = 0.45) -> float:
    """dt with vp_max*dt*sqrt(sum 1/d^2) = courant (the reference drivers' Courant
    number, e.g. test3d_homegenous.py:35)."""
    d = np.atleast_1d(np.asarray(d, dtype=np.float64))
    return float(courant / (vp_max * np.sqrt(np.sum(1.0 / d ** 2))))


def layered_2d_with_anomalies(shape, d: float, seed: int = 0, n_layers: int = 6, vp_top: float = 2000.0,
                              vp_bottom: float = 4500.0, amplitude: float = 0.10, radius: float = 150.0):
    """2-D (nx, nz) layered background and the same with two Gaussian
    anomalies, +amplitude and -amplitude in vp and vs, of e-folding radius
    ``radius`` metres, at seeded positions in the middle third of the depth.
    Interfaces are gently dipping, layer vp rises linearly with depth,
    vs = vp / 1.73, Gardner density (unchanged by the anomalies).
    Returns ((vp0, vs0, rho), (vp1, vs1, rho)) as float32 (nx, nz) arrays."""
    rng = np.random.default_rng(seed)
    nx, nz = shape
    x = np.arange(nx, dtype=np.float64) * d
    z = np.arange(nz, dtype=np.float64) * d
    depth = nz * d
    zbar = np.linspace(0.12, 0.88, n_layers - 1) * depth
    layer = np.zeros(shape, dtype=np.int16)
    for m in range(n_layers - 1):
        dip = rng.uniform(-0.05, 0.05)
        surf = zbar[m] + dip * (x - x.mean())
        layer += (z[None, :] >= surf[:, None]).astype(np.int16)
    vlayers = np.linspace(vp_top, vp_bottom, n_layers)
    vp0 = vlayers[layer]
    vs0 = vp0 / 1.73
    rho = 1741.0 * (vp0 / 1000.0) ** 0.25
    bump = np.zeros(shape)
    for sign in (1.0, -1.0):
        cx = rng.uniform(0.25, 0.75) * nx * d
        cz = rng.uniform(1.0 / 3.0, 2.0 / 3.0) * depth
        r2 = (x[:, None] - cx) ** 2 + (z[None, :] - cz) ** 2
        bump += sign * np.exp(-r2 / radius ** 2)
    vp1 = vp0 * (1.0 + amplitude * bump)
    vs1 = vs0 * (1.0 + amplitude * bump)
    f = lambda a: a.astype(np.float32)  # noqa: E731
    return (f(vp0), f(vs0), f(rho)), (f(vp1), f(vs1), f(rho))
