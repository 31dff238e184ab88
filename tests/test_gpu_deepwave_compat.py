"""The deepwave-style front end on the GPU: its receiver traces equal a direct
run of the 2-D modelling on the padded model, its lamb/mu/buoyancy gradients
equal the chain rule applied to the adjoint's vp/vs/rho gradients, and the
Compare_with_deepwave.py call shape runs."""
import numpy as np
import pytest
import torch

from seiswave_taichi_amd import deepwave_compat as D
from seiswave_taichi_amd.coeffs import ricker2
from tests.cases import layered_model

pytestmark = pytest.mark.gpu


def _setup(ny=40, nx=52, nt=150, dt=6e-4, seed=30):
    vs, vp, rho = layered_model((nx, ny), seed=seed)          # (x, z) order
    vp_t, vs_t, rho_t = (torch.from_numpy(a.T.copy()).double() for a in (vp, vs, rho))  # [ny, nx]
    lamb, mu, b = D.vpvsrho_to_lambmubuoyancy(vp_t, vs_t, rho_t)
    amps = torch.from_numpy(1e6 * ricker2(np.arange(nt), dt, 1.5 / 25.0, 25.0, 1.0)).reshape(1, 1, -1)
    src = torch.tensor([[[12, 20]]])                           # (y, x)
    rec = torch.tensor([[[3, x] for x in range(6, 46, 5)]])
    return lamb.float(), mu.float(), b.float(), amps, src, rec, dt


def test_front_end_matches_direct_modelling():
    from seiswave_taichi_amd.adjoint import Elastic2DModeling
    lamb, mu, b, amps, src, rec, dt = _setup()
    out = D.elastic(lamb, mu, b, 10.0, dt, source_amplitudes_y=amps, source_locations_y=src,
                    receiver_locations_y=rec, receiver_locations_x=rec, accuracy=4, pml_width=10, pml_freq=25.0)
    assert len(out) == 16 and out[14].shape == (1, rec.shape[1], amps.shape[2])
    vp, vs, rho = D.lambmubuoyancy_to_vpvsrho(lamb.cuda(), mu.cuda(), b.cuda())
    pad = 2 + 10
    pv = [torch.nn.functional.pad(t.t()[None, None], (pad,) * 4, mode="replicate")[0, 0] for t in (vp, vs, rho)]
    cells = [(int(x) + pad, int(y) + pad) for y, x in rec[0].tolist()]
    pml = {"vp_max": float(vp.max()), "pml_x_thick": 100.0, "pml_z_thick": 100.0, "Rcoef": 1e-3, "theta": 1,
           "alpha_max_pml": np.pi * 25.0, "kmax_pml": 1.0}
    m = Elastic2DModeling(tuple(pv[0].shape), 10.0, 10.0, dt, amps.shape[2], 2, pml, [(20 + pad, 12 + pad)],
                          cells + cells, components=("vz", "vx"), source="force_vz",
                          wavelet=amps[0, 0].numpy().astype(np.float32))
    d, _ = m.model(*pv)
    n = rec.shape[1]
    np.testing.assert_array_equal(out[14][0].detach().cpu().numpy(), d[0, :, 0, :n].t().cpu().numpy())
    np.testing.assert_array_equal(out[15][0].detach().cpu().numpy(), d[0, :, 1, n:].t().cpu().numpy())
    assert float(out[14].abs().max()) > 0


def test_front_end_gradients_follow_the_chain_rule():
    from seiswave_taichi_amd.adjoint import Elastic2DModeling, seismograms
    lamb, mu, b, amps, src, rec, dt = _setup(seed=31)
    params = [t.cuda().requires_grad_(True) for t in (lamb, mu, b)]
    out = D.elastic(*params, 10.0, dt, source_amplitudes_y=amps, source_locations_y=src,
                    receiver_locations_y=rec, accuracy=4, pml_width=10, pml_freq=25.0)
    target = 0.8 * out[14].detach()
    (0.5 * torch.sum((out[14] - target) ** 2)).backward()
    g_l, g_m, g_b = (p.grad.double().cpu().numpy() for p in params)
    # the same misfit through vp, vs, rho directly
    with torch.no_grad():
        vp, vs, rho = D.lambmubuoyancy_to_vpvsrho(*(p.detach() for p in params))
    vv = [t.clone().requires_grad_(True) for t in (vp, vs, rho)]
    pad = 12
    pv = [torch.nn.functional.pad(t.t()[None, None], (pad,) * 4, mode="replicate")[0, 0] for t in vv]
    cells = [(int(x) + pad, int(y) + pad) for y, x in rec[0].tolist()]
    pml = {"vp_max": float(vp.max()), "pml_x_thick": 100.0, "pml_z_thick": 100.0, "Rcoef": 1e-3, "theta": 1,
           "alpha_max_pml": np.pi * 25.0, "kmax_pml": 1.0}
    m = Elastic2DModeling(tuple(pv[0].shape), 10.0, 10.0, dt, amps.shape[2], 2, pml, [(20 + pad, 12 + pad)], cells,
                          components=("vz", "vx"), source="force_vz", wavelet=amps[0, 0].numpy().astype(np.float32))
    d = seismograms(*pv, m)[0, :, 0, :].t()[None]
    (0.5 * torch.sum((d - target) ** 2)).backward()
    gvp, gvs, grho = (t.grad.double().cpu().numpy() for t in vv)
    L, M, B = (p.detach().double().cpu().numpy() for p in params)
    VP, VS = vp.double().cpu().numpy(), vs.double().cpu().numpy()
    want_l = gvp * B / (2 * VP)
    want_m = gvp * B / VP + gvs * B / (2 * VS)
    want_b = gvp * (L + 2 * M) / (2 * VP) + gvs * M / (2 * VS) - grho / B ** 2
    inner = (slice(3, -3), slice(3, -3))
    for got, want in ((g_l, want_l), (g_m, want_m), (g_b, want_b)):
        assert np.linalg.norm(got[inner] - want[inner]) <= 1e-4 * np.linalg.norm(want[inner])
    assert np.abs(g_b).max() > 0


def test_compare_with_deepwave_call_shape():
    """Compare_with_deepwave.py:16-50 scaled to 128 x 128 (the reference runs
    512 x 512, 2000 steps): one y force, one y receiver, out[14][0][0]."""
    n, nt, dt, freq = 128, 400, 1e-4, 60.0
    vp = torch.ones(n, n) * 2500
    vs = torch.ones(n, n) * 1500
    rho = torch.ones(n, n) * 2200
    lamb, mu, b = D.vpvsrho_to_lambmubuoyancy(vp, vs, rho)
    amps = torch.from_numpy(ricker2(np.arange(nt), dt, 1.5 / freq, freq, 1.0)).reshape(1, 1, -1)
    out = D.elastic(lamb, mu, b, 1, dt, source_amplitudes_y=amps, source_locations_y=torch.tensor([[[40, 40]]]),
                    receiver_locations_y=torch.tensor([[[60, 60]]]), pml_freq=freq, accuracy=4)
    trace = out[14][0][0].detach().cpu().numpy()
    assert trace.shape == (nt,) and np.all(np.isfinite(trace))
    # P arrives after r/vp = 28.3 m / 2500 m/s = 11.3 ms = 113 steps (plus the wavelet delay)
    first = int(np.argmax(np.abs(trace) > 1e-3 * np.abs(trace).max()))
    assert 60 <= first <= 250, first


def test_front_end_traces_match_oracle():
    """deepwave_compat.elastic's traces (out[14]: vz at the y receivers,
    out[15]: vx at the x receivers) against the f32 oracle's 2-D VS run on the
    same padded model (the model's edge cells repeated over the stencil and
    CPML cells, R = 1e-3, alpha_max = pi*pml_freq, kmax = 1) with the same vz
    force: bit for bit."""
    from oracle import oracle as O
    lamb, mu, b, amps, src, rec, dt = _setup(seed=32)
    out = D.elastic(lamb, mu, b, 10.0, dt, source_amplitudes_y=amps, source_locations_y=src,
                    receiver_locations_y=rec, receiver_locations_x=rec, accuracy=4, pml_width=10, pml_freq=25.0)
    vp, vs, rho = (t.cpu().numpy() for t in D.lambmubuoyancy_to_vpvsrho(lamb.cuda(), mu.cuda(), b.cuda()))
    pad = 2 + 10
    pv = [np.ascontiguousarray(np.pad(a.T, pad, mode="edge")) for a in (vp, vs, rho)]  # (x, z)
    cells = np.array([(int(x) + pad, int(y) + pad) for y, x in rec[0].tolist()])
    ix, iz = 20 + pad, 12 + pad
    pml = {"vp_max": float(vp.max()), "pml_x_thick": 100.0, "pml_z_thick": 100.0, "Rcoef": 1e-3, "theta": 1,
           "alpha_max_pml": np.pi * 25.0, "kmax_pml": 1.0}
    nt = amps.shape[2]
    o = O.Oracle2D(pv[1], pv[0], pv[2], 10.0, 10.0, dt, ix, iz, 0, 0, nt, 2, freq=25.0)
    o.set_pml([True] * 4, pml)
    w = amps[0, 0].numpy().astype(np.float32)
    tz, tx = np.zeros((len(cells), nt), np.float32), np.zeros((len(cells), nt), np.float32)
    for it in range(nt):
        o.vs_step(it, float(w[it]))
        tz[:, it] = o.fields["vz"][cells[:, 0], cells[:, 1]]
        tx[:, it] = o.fields["vx"][cells[:, 0], cells[:, 1]]
    assert np.abs(tz).max() > 0
    np.testing.assert_array_equal(out[14][0].detach().cpu().numpy(), tz)
    np.testing.assert_array_equal(out[15][0].detach().cpu().numpy(), tx)
