"""A deepwave-style front end to the 2-D solver and its hand-written adjoint.

Compare_with_deepwave.py:16-45 runs ``deepwave.elastic`` on a 512 x 512 model
with ``vpvsrho_to_lambmubuoyancy``, a body force in y and a y receiver, and
reads ``out[14]``.  This module takes the same arguments and conventions, so
such a script, or a gradient cross-check against deepwave on a machine that
has it, runs on this package:

* models are [ny, nx] tensors, y (the first dimension) the depth axis; here y
  is the solver's z and the model is handed over in its (x, z) order;
* ``lamb, mu, buoyancy`` map to (vp, vs, rho) with ``rho = 1/buoyancy``,
  ``vs = sqrt(mu/rho)``, ``vp = sqrt((lamb + 2 mu)/rho)``, through torch, so
  gradients with respect to lamb, mu and buoyancy come from the adjoint's
  (vp, vs, rho) gradient by the chain rule;
* body forces ``source_amplitudes_y/x`` [shot, source, nt] at
  ``source_locations_y/x`` [shot, source, 2] ((y, x) cell indices): the force
  adds ``dt * amplitude * buoyancy`` to vz (y) or vx (x) after the velocity
  sweep; several sources or both components are run one at a time and summed
  (the step is linear in its sources);
* ``receiver_locations_y/x`` [shot, receiver, 2] record vz / vx after every
  step into ``receiver_amplitudes_y/x`` [shot, receiver, nt];
* ``accuracy`` is deepwave's order of accuracy (2, 4, ..., 16): the solver's
  staggered stencil has half-length accuracy/2;
* ``pml_width`` cells of CPML are added around the model on every side (the
  model's edge cells repeated outwards) behind ``accuracy/2`` stencil cells;
  ``pml_freq`` sets alpha_max = pi * pml_freq (default 25 Hz), R = 1e-3,
  kappa = 1; ``max_vel`` (default max vp) sets the damping.

Not deepwave: the grid is the VS staggering of src/Elastic2D_CPML.py:127-200
(vx at integer, vz at half cells) and the CPML is the reference's; dt must
satisfy the Courant limit (deepwave resamples in time instead); the returned
tuple has deepwave's 16 positions, with the receiver amplitudes filled and the
final wavefields and memory variables None; pressure receivers are refused;
gradients flow to the model, not to the source amplitudes.
"""
from __future__ import annotations

import numpy as np
import torch

from .coeffs import diff_coff

N_OUTPUTS = 16
OUT_RECEIVERS_Y, OUT_RECEIVERS_X = 14, 15


def vpvsrho_to_lambmubuoyancy(vp, vs, rho, eps: float = 1e-10):
    """(lamb, mu, buoyancy) = (rho (vp^2 - 2 vs^2), rho vs^2, 1/rho); buoyancy 0 where rho < eps."""
    vp, vs, rho = (torch.as_tensor(a) for a in (vp, vs, rho))
    mu = rho * vs ** 2
    lamb = rho * vp ** 2 - 2 * mu
    buoyancy = torch.where(rho.abs() < eps, torch.zeros_like(rho), 1.0 / torch.where(rho.abs() < eps,
                                                                                      torch.ones_like(rho), rho))
    return lamb, mu, buoyancy


def lambmubuoyancy_to_vpvsrho(lamb, mu, buoyancy):
    """(vp, vs, rho) from (lamb, mu, buoyancy); buoyancy must be positive."""
    lamb, mu, buoyancy = (torch.as_tensor(a) for a in (lamb, mu, buoyancy))
    if torch.any(buoyancy <= 0):
        raise ValueError("buoyancy must be positive everywhere")
    rho = 1.0 / buoyancy
    vs = torch.sqrt(torch.clamp(mu * buoyancy, min=0.0))
    vp = torch.sqrt((lamb + 2 * mu) * buoyancy)
    return vp, vs, rho


def _spacing(grid_spacing):
    g = np.broadcast_to(np.asarray(grid_spacing, dtype=np.float64), (2,))
    return float(g[0]), float(g[1])  # dy (depth), dx


def _locations(loc, nshot, name):
    if loc is None:
        return None
    a = torch.as_tensor(loc).detach().cpu().numpy().astype(np.int64)
    if a.ndim != 3 or a.shape[0] != nshot or a.shape[2] != 2:
        raise ValueError(f"{name} must be [n_shots, n, 2] (y, x) indices")
    return a


def elastic(lamb, mu, buoyancy, grid_spacing, dt: float, source_amplitudes_y=None, source_amplitudes_x=None,
            source_locations_y=None, source_locations_x=None, receiver_locations_y=None,
            receiver_locations_x=None, receiver_locations_p=None, accuracy: int = 4, pml_width: int = 20,
            pml_freq: float | None = None, max_vel: float | None = None, checkpoint: int = 50):
    """Forward modelling in deepwave's calling convention (see the module
    docstring); returns a 16-tuple with receiver_amplitudes_y at position 14
    and receiver_amplitudes_x at 15, differentiable with respect to lamb, mu
    and buoyancy."""
    from .adjoint import Elastic2DModeling, seismograms

    if receiver_locations_p is not None:
        raise NotImplementedError("pressure receivers are not supported")
    if accuracy % 2 or not 2 <= accuracy <= 16:
        raise ValueError("accuracy must be an even order of accuracy in 2..16")
    order = accuracy // 2
    dy, dx = _spacing(grid_spacing)
    dev = torch.device("cuda", torch.cuda.current_device())
    lamb, mu, buoyancy = (torch.as_tensor(a).to(dev, torch.float32) for a in (lamb, mu, buoyancy))
    if lamb.dim() != 2 or lamb.shape != mu.shape or lamb.shape != buoyancy.shape:
        raise ValueError("lamb, mu, buoyancy must be [ny, nx] tensors of one shape")
    ny, nx = lamb.shape
    vp, vs, rho = lambmubuoyancy_to_vpvsrho(lamb, mu, buoyancy)
    srcs = [(a, _locations(l, 0 if a is None else len(a), n), comp)
            for a, l, n, comp in ((source_amplitudes_y, source_locations_y, "source_locations_y", "force_vz"),
                                  (source_amplitudes_x, source_locations_x, "source_locations_x", "force_vx"))
            if a is not None]
    if not srcs:
        raise ValueError("no source amplitudes given")
    nshot, nt = srcs[0][0].shape[0], srcs[0][0].shape[2]
    for a, loc, _ in srcs:
        if a.shape[0] != nshot or a.shape[2] != nt or loc is None or loc.shape[1] != a.shape[1]:
            raise ValueError("source amplitudes [shot, source, nt] and locations [shot, source, 2] disagree")
    rec_y = _locations(receiver_locations_y, nshot, "receiver_locations_y")
    rec_x = _locations(receiver_locations_x, nshot, "receiver_locations_x")
    vmax = float(vp.detach().max()) if max_vel is None else float(max_vel)
    courant = vmax * dt * np.sqrt(1.0 / dx ** 2 + 1.0 / dy ** 2) * float(np.sum(np.abs(diff_coff(order))))
    if courant >= 1.0:
        raise ValueError(f"dt = {dt} exceeds the Courant limit of this grid (number {courant:.3f} >= 1); "
                         "use a smaller dt")
    # reference (x, z) order, padded by the stencil cells and the CPML on every side
    pad = order + int(pml_width)

    def padded(t):
        return torch.nn.functional.pad(t.t()[None, None], (pad, pad, pad, pad), mode="replicate")[0, 0]

    pvp, pvs, prho = padded(vp), padded(vs), padded(rho)
    shape = (nx + 2 * pad, ny + 2 * pad)
    f = 25.0 if pml_freq is None else float(pml_freq)
    pml = {"vp_max": vmax, "pml_x_thick": pml_width * dx, "pml_z_thick": pml_width * dy, "Rcoef": 1e-3,
           "theta": 1, "alpha_max_pml": np.pi * f, "kmax_pml": 1.0}

    def cells(loc_row):  # (y, x) model indices -> (x, z) padded grid cells
        return [(int(x) + pad, int(y) + pad) for y, x in loc_row]

    ys, xs = [], []
    for shot in range(nshot):
        recs = ([] if rec_y is None else cells(rec_y[shot])) + ([] if rec_x is None else cells(rec_x[shot]))
        ny_rec = 0 if rec_y is None else rec_y.shape[1]
        total = None
        for amps, loc, kind in srcs:
            for k in range(amps.shape[1]):
                w = torch.as_tensor(amps[shot, k]).detach().cpu().numpy().astype(np.float32)
                m = Elastic2DModeling(shape, dx, dy, dt, nt, order, pml, [cells(loc[shot])[k]], recs,
                                      freq=f, components=("vz", "vx"), source=kind, wavelet=w,
                                      checkpoint=checkpoint)
                d = seismograms(pvp, pvs, prho, m)[0]  # [nt, (vz, vx), receiver]
                total = d if total is None else total + d
        ys.append(total[:, 0, :ny_rec].t())
        xs.append(total[:, 1, ny_rec:].t())
    out = [None] * N_OUTPUTS
    out[OUT_RECEIVERS_Y] = None if rec_y is None else torch.stack(ys)
    out[OUT_RECEIVERS_X] = None if rec_x is None else torch.stack(xs)
    return tuple(out)
