"""Shared test cases: small seeded models, PML dicts and oracle/product runs.

The reference's drivers set the PML dict as in test3d_homegenous.py:44-58
(thickness = NPoint*d, Rcoef 1e-6, theta 1, alpha_max = pi*f, kmax = vs_max/(5 d f)).
"""
import numpy as np


def pml_dict(vp_max, npml, d, freq, vs_max, dims=3):
    p = {"vp_max": vp_max, "pml_x_thick": npml * d[0], "Rcoef": 1e-6, "theta": 1,
         "alpha_max_pml": np.pi * freq, "kmax_pml": vs_max / (5 * d[0] * freq)}
    if dims == 3:
        p["pml_y_thick"] = npml * d[1]
        p["pml_z_thick"] = npml * d[2]
    else:
        p["pml_z_thick"] = npml * d[1]
    return p


def layered_model(shape, seed=0, vp0=2500.0, vp1=4500.0, water=False):
    """Seeded heterogeneous model: layered vp rising with the last axis plus a
    random perturbation, vs = vp/1.73 (0 in a top layer when ``water``), Gardner rho."""
    rng = np.random.default_rng(seed)
    nz = shape[-1]
    depth = np.linspace(0, 1, nz)
    vp = np.broadcast_to(vp0 + (vp1 - vp0) * depth, shape).astype(np.float64).copy()
    vp *= 1 + 0.05 * rng.standard_normal(shape)
    vs = vp / 1.73
    if water:
        vs[..., : max(2, nz // 8)] = 0.0
    rho = 1741.0 * (vp / 1000.0) ** 0.25
    return vs.astype(np.float32), vp.astype(np.float32), rho.astype(np.float32)


def homogeneous(shape, vp=3500.0, vs=2000.0, rho=2200.0):
    return (np.full(shape, vs, np.float32), np.full(shape, vp, np.float32), np.full(shape, rho, np.float32))


def unit_mt(seed=0):
    """Six seeded normals (xx, yy, zz, xy, xz, yz) scaled to Frobenius norm 1
    (off-diagonals counted twice)."""
    m = np.random.default_rng(seed).standard_normal(6)
    m /= np.sqrt(m[0] ** 2 + m[1] ** 2 + m[2] ** 2 + 2 * (m[3] ** 2 + m[4] ** 2 + m[5] ** 2))
    return m.astype(np.float32)


def rel_l2(a, b):
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    den = np.linalg.norm(b)
    return float(np.linalg.norm(a - b) / den) if den > 0 else float(np.linalg.norm(a - b))


def per_trace_rel_l2(a, b):
    """Relative L2 per trace (traces along the last axis of (nt, ntraces))."""
    a, b = np.asarray(a, np.float64), np.asarray(b, np.float64)
    a = a.reshape(a.shape[0], -1)
    b = b.reshape(b.shape[0], -1)
    num = np.linalg.norm(a - b, axis=0)
    den = np.linalg.norm(b, axis=0)
    return np.where(den > 0, num / np.maximum(den, 1e-300), num)
