"""The benchmark configurations at their full sizes on the GPU.

* 600x600x400 layered model with eight folded interfaces, a water layer
  (vs = 0), Gardner density, CFL 0.45 and 200x200 receivers (the five-launch
  sweep plan): bit for bit against the oracle over the first steps of a
  moment source placed so that every launch, the water layer and the
  receivers carry signal; and the full 3000-step run, checked through the
  linearity of its 40,000 traces in the source.
* 200^3 homogeneous case through the split-field class with the seeded
  unit moment tensor: bit for bit against the oracle (fields, all 24 split
  components, receivers) with the source near a corner so the split region
  carries signal; and, with the source at the centre, equal to the Tensor
  (CPML) class before the wave reaches the boundary.
* 800x300 adjoint gradient (16 explosive shots, 400 receivers, 2500 steps,
  checkpoint every 50): one shot's seismograms bit for bit against the f32
  oracle, the shot batch against shot-by-shot runs, the transpose identity
  of a 50-step segment, and the gradient against directional central
  differences along smooth perturbations of vp, vs and rho.
"""
import math

import numpy as np
import pytest
import torch

from oracle import oracle as O
from seiswave_taichi_amd import _lib as L
from seiswave_taichi_amd.coeffs import ricker2
from tests.cases import homogeneous, pml_dict, unit_mt

pytestmark = pytest.mark.gpu


def _mt3(m):
    mt3 = np.zeros((3, 3), np.float32)
    mt3[0, 0], mt3[1, 1], mt3[2, 2], mt3[0, 1], mt3[0, 2], mt3[1, 2] = m
    return mt3


# ------------------------------------------------------ 600 x 600 x 400 folded
def _config3():
    from seiswave_taichi_amd.models import layered_folded_model
    shape, d, f = (600, 600, 400), 12.5, 10.0
    vs, vp, rho = layered_folded_model(shape, d, seed=0)
    dt = 0.45 * d / (float(vp.max()) * np.sqrt(3.0))
    lin = np.round(np.linspace(3, shape[0] - 4, 200)).astype(np.int64)
    rx, ry = (a.ravel() for a in np.meshgrid(lin, lin, indexing="ij"))
    rz = np.full_like(rx, 2 + 5)
    p = pml_dict(float(vp.max()), 20, (d, d, d), f, float(vs.max()))
    return shape, d, f, vs, vp, rho, dt, (rx, ry, rz), p


def test_config3_folded_model_matches_oracle():
    """600x600x400 folded model with its water layer: 8 steps bit for bit
    against the oracle; the source sits in the water near a corner, so the
    x-only bulk launch, the y-slab and z-slab launches, the water layer and
    the 200x200 receiver grid all carry signal."""
    from seiswave_taichi_amd.Elastic3D_SSG_Tensor import ElasticWAVE
    shape, d, f, vs, vp, rho, dt, (rx, ry, rz), p = _config3()
    assert float(vs[:, :, :10].max()) == 0.0  # a water layer on top
    steps, src, m = 8, (24, 25, 12), unit_mt(0)
    amps = 1e6 * ricker2(np.arange(steps) + 20, dt, 0.03, f, 1.0)
    w = ElasticWAVE(vs, vp, rho, d, d, d, dt, *src, rx[None], ry[None], rz[None], _mt3(m), 1.0, steps, 2, f,
                    source={"kind": "moment", "cell": src, "mt": m, "wavelet": amps})
    w.SetADEPML3D([True] * 6, p)
    assert L.launch_plan(w._e.descriptor()) == [1, 3, 3, 7, 7]
    w.run(steps)
    o = O.Oracle3D("tensor", vs, vp, rho, d, d, d, dt, src, np.stack([rx, ry, rz], 1), steps, 2, freq=f, mt=m)
    o.set_pml([True] * 6, p)
    o.run(steps, amps=amps)
    for name in o.FIELDS:
        np.testing.assert_array_equal(getattr(w, name).to_numpy(), o.fields[name], err_msg=name)
    np.testing.assert_array_equal(w.data.to_numpy().reshape(o.data.shape), o.data)
    vx = o.fields["vx"]
    lo = 23  # first index outside the 20-cell CPML slabs (2 + 20 + 1)
    assert np.abs(vx[lo:, lo:, lo:]).max() > 0, "bulk launch without signal"
    assert np.abs(vx[lo:, :lo, lo:]).max() > 0 and np.abs(vx[:, :, :lo]).max() > 0
    assert np.abs(o.fields["sxx"][vs == 0]).max() > 0, "no signal in the water layer"
    assert np.count_nonzero(o.data[-1]) > 10
    del w, o
    torch.cuda.empty_cache()


def test_config3_full_run_is_linear_in_the_source():
    """The full configuration: 3000 steps of the moment source at the centre
    (300, 300, 200) with 40,000 receivers.  The traces are finite and record
    the arrivals, and doubling the source amplitude doubles every trace to
    within the 1e-5 parity tolerance (measured 4.4e-6).  Not exactly: f32
    rounding is scale-invariant except where a result is subnormal, and the
    far precursor of the stencil light cone, which decays faster than any
    exponential with distance, passes through subnormal values whatever the
    source scale; one rounding there decorrelates the two runs at the level
    of an ulp (the CPU oracle does the same; f64 runs are exactly linear)."""
    from seiswave_taichi_amd.Elastic3D_SSG_Tensor import ElasticWAVE
    shape, d, f, vs, vp, rho, dt, (rx, ry, rz), p = _config3()
    nt, src, m = 3000, (300, 300, 200), unit_mt(0)
    out = []
    for scale in (1.0, 2.0):
        amps = scale * ricker2(np.arange(nt), dt, 1.5 / f, f, 1.0)
        w = ElasticWAVE(vs, vp, rho, d, d, d, dt, *src, rx[None], ry[None], rz[None], _mt3(m), 1.0, nt, 2, f,
                        source={"kind": "moment", "cell": src, "mt": m, "wavelet": amps})
        w.SetADEPML3D([True] * 6, p)
        w.run(nt)
        out.append(w.data.to_numpy().reshape(nt, -1).astype(np.float64))
        del w
        torch.cuda.empty_cache()
    a, b = out
    assert np.all(np.isfinite(a)) and np.abs(a).max() > 0
    live = np.linalg.norm(a, axis=0) > 1e-3 * np.linalg.norm(a, axis=0).max()
    assert live.sum() > 1000  # the arrivals reach a large part of the surface grid
    err = np.linalg.norm(b - 2 * a) / np.linalg.norm(2 * a)
    print("linearity of the 3000-step traces:", err)
    assert err < 1e-5, err


# -------------------------------------------------------------- 200^3 SPML
def _config2(src, nt, surface=(True,) * 6):
    n, d, f = 200, (10.0, 10.0, 10.0), 10.0
    vs, vp, rho = homogeneous((n, n, n))
    lin = np.round(np.linspace(3, n - 4, 50)).astype(np.int64)
    rx, ry = (a.ravel() for a in np.meshgrid(lin, lin, indexing="ij"))
    rec = np.stack([rx, ry, np.full_like(rx, 7)], 1)
    p = {"vp_max": 3500.0, "pml_x_thick": 200.0, "pml_y_thick": 200.0, "pml_z_thick": 200.0, "Rcoef": 1e-6,
         "theta": 1, "alpha_max_pml": np.pi * f, "kmax_pml": 4.0}
    return vs, vp, rho, d, 1e-3, f, rec, p


def test_config2_spml_moment_source_matches_oracle():
    """200^3 split-field class with the seeded unit moment tensor (an
    extension of src/Elastic3D_SSG_SPML.py, whose own source is a vz force):
    12 steps bit for bit against the oracle, every field, all 24 split
    components and the 2,500 receivers, with the source near a corner so the
    split region carries signal."""
    from seiswave_taichi_amd.Elastic3D_SSG_SPML import ElasticWAVE
    from seiswave_taichi_amd._classes3d import SPLIT_NAMES
    steps, src, m = 12, (30, 32, 28), unit_mt(0)
    vs, vp, rho, d, dt, f, rec, p = _config2(src, steps)
    amps = 1e6 * ricker2(np.arange(steps), dt, 1.5 / f, f, 1.0)
    w = ElasticWAVE(vs, vp, rho, *d, dt, *src, 0, 0, 0, steps, 2, f,
                    source={"kind": "moment", "cell": src, "mt": m, "wavelet": amps}, receivers=rec)
    w.SetADEPML3D([True] * 6, p)
    w.run(steps)
    o = O.Oracle3D("spml", vs, vp, rho, *d, dt, src, rec, steps, 2, freq=f, mt=m)
    o.set_pml([True] * 6, p)
    for it in range(steps):
        o.step_moment(it, float(np.float32(amps[it])))
    for name in o.FIELDS:
        np.testing.assert_array_equal(getattr(w, name).to_numpy(), o.fields[name], err_msg=name)
    for i, name in enumerate(SPLIT_NAMES):
        np.testing.assert_array_equal(w.split_field(name), o.split[i], err_msg=name)
    np.testing.assert_array_equal(w.data.to_numpy().reshape(o.data.shape), o.data)
    assert sum(np.count_nonzero(q) for q in o.split) > 10000
    del w, o
    torch.cuda.empty_cache()


def test_config2_spml_equals_tensor_before_the_boundary():
    """200^3, seeded unit moment tensor at the centre, 150 steps (the P wave
    travels 52 cells, the split region starts 78 cells away): the split-field
    class and the Tensor (CPML) class agree to f32 rounding (their interior
    arithmetic differs only in the order of the velocity sum and in
    multiplying by 1/rho rather than dividing by rho)."""
    from seiswave_taichi_amd.Elastic3D_SSG_SPML import ElasticWAVE as SPML
    from seiswave_taichi_amd.Elastic3D_SSG_Tensor import ElasticWAVE as Tensor
    steps, src, m = 150, (100, 100, 100), unit_mt(0)
    vs, vp, rho, d, dt, f, rec, p = _config2(src, steps)
    amps = ricker2(np.arange(steps), dt, 1.5 / f, f, 1.0)
    source = {"kind": "moment", "cell": src, "mt": m, "wavelet": amps}
    a = SPML(vs, vp, rho, *d, dt, *src, 0, 0, 0, steps, 2, f, source=source, receivers=rec)
    b = Tensor(vs, vp, rho, *d, dt, *src, rec[:, 0][None], rec[:, 1][None], rec[:, 2][None], _mt3(m), 1.0,
               steps, 2, f, source=source)
    for s in (a, b):
        s.SetADEPML3D([True] * 6, p)
        s.run(steps)
    for name in O.Oracle3D.FIELDS:
        x, y = getattr(a, name).to_numpy().astype(np.float64), getattr(b, name).to_numpy().astype(np.float64)
        assert np.abs(y).max() > 0
        err = np.linalg.norm(x - y) / np.linalg.norm(y)
        assert err < 1e-5, (name, err)
    del a, b
    torch.cuda.empty_cache()


# ------------------------------------------------------ 800 x 300 gradient
def _config5(batch=True):
    from seiswave_taichi_amd.adjoint import Elastic2DModeling
    from seiswave_taichi_amd.models import cfl_dt, layered_2d_with_anomalies
    nx, nz, d, nt, npml, f = 800, 300, 10.0, 2500, 20, 10.0
    (vp0, vs0, rho), (vp1, vs1, _) = layered_2d_with_anomalies((nx, nz), d, seed=0)
    dt = cfl_dt(float(vp1.max()), (d, d))
    pml = {"vp_max": float(vp1.max()), "pml_x_thick": npml * d, "pml_z_thick": npml * d, "Rcoef": 1e-6,
           "theta": 1, "alpha_max_pml": np.pi * f, "kmax_pml": 1.0}
    top = 2 + npml + 2
    shots = [(int(x), top) for x in np.round(np.linspace(40, nx - 41, 16))]
    rec = np.stack([np.round(np.linspace(30, nx - 31, 400)).astype(np.int64), np.full(400, top)], 1)
    wav = 1e8 * ricker2(np.arange(nt), dt, 1.5 / f, f, 1.0)
    m = Elastic2DModeling((nx, nz), d, d, dt, nt, 2, pml, shots, rec, freq=f, checkpoint=50, wavelet=wav,
                          batch=batch)
    return m, (vp0, vs0, rho), (vp1, vs1, rho), pml, shots, rec


def _t(a):
    return torch.from_numpy(np.ascontiguousarray(a))


def test_config5_seismograms_match_oracle():
    """One shot of the 800x300 configuration over all 2500 steps: the batched
    GPU seismograms (vx, vz at 400 receivers) equal the f32 oracle's with the
    same explosive pre-step injection, bit for bit."""
    m, (vp0, vs0, rho), _, pml, shots, rec = _config5()
    data, _ = m.model(_t(vp0), _t(vs0), _t(rho))
    got = data[5].cpu().numpy()
    ix, iz = shots[5]
    o = O.Oracle2D(vs0, vp0, rho, m.dx, m.dz, m.dt, ix, iz, 0, 0, m.nt, 2, freq=10.0)
    o.set_pml([True] * 4, pml)
    amps = m.amp.cpu().numpy()
    ref = np.zeros_like(got)
    for it in range(m.nt):
        e = np.float32(amps[it]) * np.float32(-0.70710678118654752)
        o.fields["sxx"][ix, iz] += e
        o.fields["szz"][ix, iz] += e
        o.vs_step(it, 0.0)
        ref[it, 0] = o.fields["vx"][rec[:, 0], rec[:, 1]]
        ref[it, 1] = o.fields["vz"][rec[:, 0], rec[:, 1]]
    assert np.abs(ref).max() > 0
    np.testing.assert_array_equal(got, ref)


def test_config5_batch_equals_shot_by_shot():
    """All 16 shots in each launch against one shot at a time: seismograms
    bit for bit, gradients to f32 rounding (the per-shot accumulators are
    summed in a fixed order at the end instead of shot after shot)."""
    from seiswave_taichi_amd.adjoint import seismograms
    res = []
    for batch in (True, False):
        m, (vp0, vs0, rho), (vp1, vs1, _), *_ = _config5(batch)
        dobs, _ = m.model(_t(vp1), _t(vs1), _t(rho))
        tv = [_t(a).cuda().requires_grad_(True) for a in (vp0, vs0, rho)]
        data = seismograms(*tv, m)
        loss = 0.5 * torch.sum((data.double() - dobs.double()) ** 2)
        loss.backward()
        res.append((data.detach().cpu().numpy(), [t.grad.double().cpu().numpy() for t in tv]))
    (d1, g1), (d2, g2) = res
    np.testing.assert_array_equal(d1, d2)
    for a, b in zip(g1, g2):
        assert np.abs(b).max() > 0
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)


def test_config5_segment_transpose_identity():
    """<y, A^50 x> = <(A^T)^50 y, x> for one 50-step checkpoint segment of
    the 800x300 configuration (CPML memory variables included), A^T applied
    as the hand-written adjoint stages in reverse order."""
    m, (vp0, vs0, rho), *_ = _config5()
    m.set_model(_t(vp0), _t(vs0), _t(rho))
    lib, st = m.lib, m.stream()
    g = torch.Generator(device=m.device).manual_seed(11)
    x = torch.randn((15, m.nz, m.nx), generator=g, device=m.device)
    y = torch.randn((15, m.nz, m.nx), generator=g, device=m.device)
    s = x.clone()
    gd = m.desc(s)
    K = 50
    for _ in range(K):
        L.check(lib.sw_ssg2d_velocity(gd, st), "velocity")
        L.check(lib.sw_ssg2d_stress(gd, st), "stress")
    adj = y.clone()
    tmp = torch.zeros((4, m.nz, m.nx), device=m.device)
    acc = torch.zeros((5, m.nz, m.nx), device=m.device)
    q = m._adj_desc(adj, tmp, acc)
    fwd = m.desc(x.clone())
    for _ in range(K):
        L.check(lib.sw_ssg2d_adjoint_stress(fwd, q, st), "adjoint stress")
        L.check(lib.sw_ssg2d_adjoint_velocity(fwd, q, st), "adjoint velocity")
    torch.cuda.synchronize()
    lhs = torch.sum(y.double() * s.double()).item()
    rhs = torch.sum(adj.double() * x.double()).item()
    assert abs(lhs - rhs) <= 1e-4 * max(abs(lhs), abs(rhs)), (lhs, rhs)


def test_config5_gradient_against_directional_differences():
    """d J / d eps of J(m + eps dm) at eps = 0, by the fourth-order central
    difference of the f32 forward (eps = +-h, +-2h, h = 0.5), against
    <grad J, dm> from the adjoint, for a smooth dm (a Gaussian bump of 400 m
    radius at 1 % of the parameter, under the receivers between the shots)
    of vp, vs and rho in turn: relative difference <= 1e-3 (measured
    6e-5, 2e-4, 2e-4).  The second-order difference is limited by its own
    truncation error here (1.7e-3 for vp at h = 1, falling as h^2)."""
    from seiswave_taichi_amd.adjoint import seismograms
    m, (vp0, vs0, rho), (vp1, vs1, _), *_ = _config5()
    dobs, _ = m.model(_t(vp1), _t(vs1), _t(rho))
    dobs = dobs.double()
    tv = [_t(a).cuda().requires_grad_(True) for a in (vp0, vs0, rho)]
    data = seismograms(*tv, m)
    loss = 0.5 * torch.sum((data.double() - dobs) ** 2)
    loss.backward()
    grads = [t.grad.double().cpu().numpy() for t in tv]
    nx, nz = vp0.shape
    x = np.arange(nx)[:, None] * 10.0
    z = np.arange(nz)[None, :] * 10.0
    bump = np.exp(-((x - 4000.0) ** 2 + (z - 1500.0) ** 2) / 400.0 ** 2)

    def misfit(params):
        dd, _ = m.model(*(_t(a.astype(np.float32)) for a in params))
        return 0.5 * float(torch.sum((dd.double() - dobs) ** 2))

    base = [vp0.astype(np.float64), vs0.astype(np.float64), rho.astype(np.float64)]
    report = {}
    for k, name in enumerate(("vp", "vs", "rho")):
        dm = 0.01 * base[k] * bump
        h = 0.5
        J = {}
        for e in (-2, -1, 1, 2):
            p = [b.copy() for b in base]
            p[k] += e * h * dm
            J[e] = misfit(p)
        fd = (-J[2] + 8 * J[1] - 8 * J[-1] + J[-2]) / (12 * h)
        ad = float(np.sum(grads[k] * dm))
        report[name] = (ad, fd, abs(ad - fd) / abs(fd))
    print(report)
    for name, (ad, fd, rel) in report.items():
        assert rel <= 1e-3, (name, ad, fd, rel)
