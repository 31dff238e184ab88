"""The hand-written 2-D adjoint: transpose identities of each stage and of a
12-step chain, and the model gradient against central finite differences of
the f64 oracle."""
import numpy as np
import pytest
import torch

from oracle import oracle as O
from seiswave_taichi_amd.coeffs import ricker2
from tests.cases import layered_model, pml_dict

pytestmark = pytest.mark.gpu


def _modeling(shape=(60, 48), nt=160, order=2, shots=((30, 14),), ck=40):
    from seiswave_taichi_amd.adjoint import Elastic2DModeling
    d = (10.0, 10.0)
    rx = np.arange(8, shape[0] - 8, 4)
    rec = np.stack([rx, np.full_like(rx, 5)], 1)
    p = pml_dict(4700.0, 8, d, 20.0, 2700.0, dims=2)
    # unit-scale seismograms (~1e-2): an fp32 adjoint of a misfit built on
    # 1e-9 velocities would push gradients into the denormal range
    wavelet = 1e8 * ricker2(np.arange(nt), 8e-4, 1.5 / 20.0, 20.0, 1.0)
    m = Elastic2DModeling(shape, *d, 8e-4, nt, order, p, shots, rec, freq=20.0, checkpoint=ck, wavelet=wavelet)
    return m, p, rec


def _rand_state(m, seed):
    g = torch.Generator(device=m.device).manual_seed(seed)
    x = torch.randn((15, m.nz, m.nx), generator=g, device=m.device, dtype=torch.float32)
    return x


@pytest.mark.parametrize("stage", ["stress", "velocity"])
def test_stage_transpose_dot_product(stage):
    """<y, S x> = <S^T y, x> for each stage map S (state incl. memory variables)."""
    from seiswave_taichi_amd import _lib as L
    m, _, _ = _modeling()
    vs, vp, rho = layered_model((m.nx, m.nz), seed=1)
    m.set_model(torch.from_numpy(vp), torch.from_numpy(vs), torch.from_numpy(rho))
    lib, st = m.lib, m.stream()
    x = _rand_state(m, 1)
    y = _rand_state(m, 2)
    sx = x.clone()
    gx = m.desc(sx)
    fn = lib.sw_ssg2d_stress if stage == "stress" else lib.sw_ssg2d_velocity
    L.check(fn(gx, st), stage)
    adj = y.clone()
    tmp = torch.zeros((4, m.nz, m.nx), device=m.device)
    acc = torch.zeros((5, m.nz, m.nx), device=m.device)
    q = m._adj_desc(adj, tmp, acc)
    fwd = m.desc(x.clone())
    afn = lib.sw_ssg2d_adjoint_stress if stage == "stress" else lib.sw_ssg2d_adjoint_velocity
    L.check(afn(fwd, q, st), "adjoint " + stage)
    torch.cuda.synchronize()
    lhs = torch.sum(y.double() * sx.double()).item()
    rhs = torch.sum(adj.double() * x.double()).item()
    assert abs(lhs - rhs) <= 1e-5 * max(abs(lhs), 1.0), (lhs, rhs)


def _oracle_misfit(vp, vs, rho, m, rec, p, dobs, shots):
    """f64 oracle forward with explosive pre-injection; returns (J, data)."""
    total = 0.0
    datas = []
    amps = m.amp.cpu().numpy().astype(np.float64)
    for s, (ix, iz) in enumerate(shots):
        o = O.Oracle2D(vs, vp, rho, m.dx, m.dz, m.dt, ix, iz, 0, 0, m.nt, m.order, freq=20.0, precision="f64")
        o.set_pml([True] * 4, p)
        data = np.zeros((m.nt, 2, len(rec)))
        for it in range(m.nt):
            e = amps[it] * (-0.70710678118654752)
            o.fields["sxx"][ix, iz] += e
            o.fields["szz"][ix, iz] += e
            o.vs_step(it, 0.0)
            data[it, 0] = o.fields["vx"][rec[:, 0], rec[:, 1]]
            data[it, 1] = o.fields["vz"][rec[:, 0], rec[:, 1]]
        datas.append(data)
        if dobs is not None:
            total += 0.5 * float(np.sum((data - dobs[s]) ** 2))
    return total, np.stack(datas)


def test_chain_transpose_dot_product():
    """<y, A^K x> = <(A^T)^K y, x> for K = 12 full steps (velocity then stress
    sweep, memory variables included), A^T applied as the adjoint stages in
    reverse order."""
    from seiswave_taichi_amd import _lib as L
    m, _, _ = _modeling()
    vs, vp, rho = layered_model((m.nx, m.nz), seed=2)
    m.set_model(torch.from_numpy(vp), torch.from_numpy(vs), torch.from_numpy(rho))
    lib, st = m.lib, m.stream()
    x = _rand_state(m, 3)
    y = _rand_state(m, 4)
    s = x.clone()
    g = m.desc(s)
    K = 12
    for _ in range(K):
        L.check(lib.sw_ssg2d_velocity(g, st), "velocity")
        L.check(lib.sw_ssg2d_stress(g, st), "stress")
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
    # f32 rounding accumulates over 24 stage applications in each direction
    assert abs(lhs - rhs) <= 1e-4 * max(abs(lhs), abs(rhs), 1.0), (lhs, rhs)


@pytest.mark.parametrize("source,components", [("explosive", ("vx", "vz")), ("force_vz", ("vx", "vz")),
                                               ("explosive", ("vz", "sxx"))])
def test_segment_matches_stage_by_stage_schedule(source, components):
    """sw_ssg2d_adjoint_segment (forward derivative sums saved by the
    recomputed sweeps; with velocity residuals each gather fused with the next
    stage's per-cell pass, with a stress residual the four passes apart)
    leaves the adjoint state and the five coefficient accumulators bit-equal
    to the per-stage calls on full forward states kept before and after every
    velocity sweep (which recompute those sums)."""
    from seiswave_taichi_amd import _lib as L
    from seiswave_taichi_amd.adjoint import Elastic2DModeling
    d, nt, K = (10.0, 10.0), 60, 25
    shape, shot = (56, 44), (27, 15)
    rx = np.arange(8, shape[0] - 8, 5)
    rec = np.stack([rx, np.full_like(rx, 5)], 1)
    p = pml_dict(4700.0, 8, d, 20.0, 2700.0, dims=2)
    wavelet = 1e8 * ricker2(np.arange(nt), 8e-4, 1.5 / 20.0, 20.0, 1.0)
    m = Elastic2DModeling(shape, *d, 8e-4, nt, 2, p, [shot], rec, freq=20.0, checkpoint=K, wavelet=wavelet,
                          source=source, batch=False, components=components)
    vs, vp, rho = layered_model(shape, seed=6)
    m.set_model(torch.from_numpy(vp), torch.from_numpy(vs), torch.from_numpy(rho))
    lib, st = m.lib, m.stream()
    force = source != "explosive"
    src = m.source(shot)
    data = torch.zeros((nt, len(m.components), m.nrec), device=m.device)
    rec_d = m.receivers(data)
    state = m.new_state()
    g = m.desc(state)
    pre, mid, ck = [], [], []
    for it in range(nt):
        if it % K == 0:
            ck.append(state.clone())
        if not force:
            L.check(lib.sw_ssg2d_inject(g, src, it, st), "inject")
        pre.append(state.clone())
        L.check(lib.sw_ssg2d_velocity(g, st), "velocity")
        if force:
            L.check(lib.sw_ssg2d_inject(g, src, it, st), "inject")
        mid.append(state.clone())
        L.check(lib.sw_ssg2d_stress(g, st), "stress")
        L.check(lib.sw_ssg2d_record(g, rec_d, it, st), "record")
    gen = torch.Generator(device=m.device).manual_seed(9)
    res = torch.randn(data.shape, generator=gen, device=m.device) * data.abs().max()
    rres = m.receivers(res)
    out = []
    for staged in (True, False):
        adj = m.new_state()
        tmp = torch.zeros((4, m.nz, m.nx), device=m.device)
        acc = torch.zeros((5, m.nz, m.nx), device=m.device)
        g_src = torch.zeros(1, device=m.device)
        q = m._adj_desc(adj, tmp, acc, g_src if force else None)
        if staged:
            for it in reversed(range(nt)):
                L.check(lib.sw_ssg2d_adjoint_inject(g, q, rres, it, st), "adjoint inject")
                L.check(lib.sw_ssg2d_adjoint_stress(m.desc(mid[it]), q, st), "adjoint stress")
                L.check(lib.sw_ssg2d_adjoint_velocity(m.desc(pre[it]), q, st), "adjoint velocity")
        else:
            work = m.new_state()
            seg = m._segment_buffer(K)
            for s0 in reversed(range(0, nt, K)):
                work.copy_(ck[s0 // K])
                L.check(lib.sw_ssg2d_adjoint_segment(m.desc(work), src, q, rres, L.fptr(seg), s0, min(K, nt - s0),
                                                     st), "adjoint segment")
        torch.cuda.synchronize()
        out.append((adj.clone(), acc.clone()))
    assert out[0][1].abs().max() > 0
    assert torch.equal(out[0][0], out[1][0]) and torch.equal(out[0][1], out[1][1])


def test_forward_modeling_matches_oracle():
    m, p, rec = _modeling()
    vs, vp, rho = layered_model((m.nx, m.nz), seed=3)
    data, _ = m.model(torch.from_numpy(vp), torch.from_numpy(vs), torch.from_numpy(rho))
    _, ref = _oracle_misfit(vp.astype(np.float64), vs.astype(np.float64), rho.astype(np.float64), m, rec, p,
                            None, m.shots)
    got = data.cpu().numpy()
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    assert err < 1e-5, err


def test_gradient_matches_finite_differences():
    """Adjoint gradient (f32, GPU) vs central differences of the f64 oracle
    misfit at 32 random cells, for vp, vs and rho: relative error <= 1e-3."""
    from seiswave_taichi_amd.adjoint import seismograms
    m, p, rec = _modeling()
    vs, vp, rho = layered_model((m.nx, m.nz), seed=4)
    # observed data: a model with a Gaussian anomaly
    ii, jj = np.meshgrid(np.arange(m.nx), np.arange(m.nz), indexing="ij")
    bump = np.exp(-((ii - 30) ** 2 + (jj - 24) ** 2) / (2 * 4.0 ** 2))
    vp_t, vs_t = (vp * (1 + 0.08 * bump)).astype(np.float32), (vs * (1 - 0.06 * bump)).astype(np.float32)
    _, dobs = _oracle_misfit(vp_t.astype(np.float64), vs_t.astype(np.float64), rho.astype(np.float64), m, rec, p,
                             None, m.shots)
    tv = [torch.from_numpy(a).to("cuda").requires_grad_(True) for a in (vp, vs, rho)]
    data = seismograms(*tv, m)
    dob = torch.from_numpy(dobs.astype(np.float32)).to("cuda")
    loss = 0.5 * torch.sum((data - dob) ** 2)
    loss.backward()
    grads = [t.grad.cpu().numpy().astype(np.float64) for t in tv]
    # 32 random cells outside the PML and at least 4 cells from the source
    # (which sits below the PML, 9 cells under the receiver line): at the
    # source the gradient is dominated by the near field, where an f32 adjoint
    # and f64 finite differences part ways by more than 1e-3
    rng = np.random.default_rng(0)
    cand = [(a, b) for a in range(12, m.nx - 12) for b in range(12, m.nz - 12)
            if max(abs(a - m.shots[0][0]), abs(b - m.shots[0][1])) >= 4]
    cells = np.array([cand[t] for t in rng.choice(len(cand), 32, replace=False)])
    base = [vp.astype(np.float64), vs.astype(np.float64), rho.astype(np.float64)]
    for k, name in enumerate(("vp", "vs", "rho")):
        fd, ad = [], []
        for (ci, cj) in cells:
            h = 1e-4 * base[k][ci, cj]
            mp = [b.copy() for b in base]
            mm = [b.copy() for b in base]
            mp[k][ci, cj] += h
            mm[k][ci, cj] -= h
            jp, _ = _oracle_misfit(mp[0], mp[1], mp[2], m, rec, p, dobs, m.shots)
            jm, _ = _oracle_misfit(mm[0], mm[1], mm[2], m, rec, p, dobs, m.shots)
            fd.append((jp - jm) / (2 * h))
            ad.append(grads[k][ci, cj])
        fd, ad = np.array(fd), np.array(ad)
        rel = np.linalg.norm(ad - fd) / np.linalg.norm(fd)
        assert rel <= 1e-3, (name, rel, fd[:4], ad[:4])


def test_shot_batch_matches_shot_by_shot():
    """Three shots in one batch: seismograms bit-equal to shot-by-shot runs;
    the gradient (accumulators summed per shot, then over shots) within f32
    rounding of the shot-by-shot one."""
    from seiswave_taichi_amd.adjoint import seismograms
    shots = ((30, 14), (18, 20), (44, 16))
    mb, p, rec = _modeling(shots=shots, ck=30)
    ms, _, _ = _modeling(shots=shots, ck=30)
    ms.batch = False
    vs, vp, rho = layered_model((mb.nx, mb.nz), seed=6)
    db, _ = mb.model(torch.from_numpy(vp), torch.from_numpy(vs), torch.from_numpy(rho))
    ds, _ = ms.model(torch.from_numpy(vp), torch.from_numpy(vs), torch.from_numpy(rho))
    np.testing.assert_array_equal(db.cpu().numpy(), ds.cpu().numpy())
    target = (0.9 * db).detach()
    grads = []
    for m in (mb, ms):
        tv = [torch.from_numpy(a).to("cuda").requires_grad_(True) for a in (vp, vs, rho)]
        loss = 0.5 * torch.sum((seismograms(*tv, m) - target) ** 2)
        loss.backward()
        grads.append([t.grad.double().cpu().numpy() for t in tv])
    for a, b in zip(*grads):
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b), np.linalg.norm(a - b) / np.linalg.norm(b)


def _oracle_misfit_force(vp, vs, rho, m, rec, p, dobs, shots, amps):
    """f64 oracle forward of the VS step with its own vz force
    (vz[p] += dt*amp/rho[p] after the velocity loops); returns (J, data)."""
    total = 0.0
    datas = []
    for s, (ix, iz) in enumerate(shots):
        o = O.Oracle2D(vs, vp, rho, m.dx, m.dz, m.dt, ix, iz, 0, 0, m.nt, m.order, freq=20.0, precision="f64")
        o.set_pml([True] * 4, p)
        data = np.zeros((m.nt, 2, len(rec)))
        for it in range(m.nt):
            o.vs_step(it, float(amps[it]))
            data[it, 0] = o.fields["vx"][rec[:, 0], rec[:, 1]]
            data[it, 1] = o.fields["vz"][rec[:, 0], rec[:, 1]]
        datas.append(data)
        if dobs is not None:
            total += 0.5 * float(np.sum((data - dobs[s]) ** 2))
    return total, np.stack(datas)


def test_force_source_gradient_matches_finite_differences():
    """A vz force (dt*amp/rho at the source, after the velocity sweep): the
    adjoint gradient against central differences of the f64 oracle at 12
    cells per parameter (<= 1e-3), and d/d rho at the source cell itself,
    where the force's own -dt*amp/rho**2 term enters (<= 1e-2)."""
    from seiswave_taichi_amd.adjoint import Elastic2DModeling, seismograms
    shape, d, nt = (60, 48), (10.0, 10.0), 160
    rx = np.arange(8, shape[0] - 8, 4)
    rec = np.stack([rx, np.full_like(rx, 5)], 1)
    p = pml_dict(4700.0, 8, d, 20.0, 2700.0, dims=2)
    shots = ((30, 14),)
    amps = 1e6 * ricker2(np.arange(nt), 8e-4, 1.5 / 20.0, 20.0, 1.0)
    m = Elastic2DModeling(shape, *d, 8e-4, nt, 2, p, shots, rec, freq=20.0, checkpoint=40, wavelet=amps,
                          source="force_vz")
    vs, vp, rho = layered_model(shape, seed=14)
    ii, jj = np.meshgrid(np.arange(shape[0]), np.arange(shape[1]), indexing="ij")
    bump = np.exp(-((ii - 30) ** 2 + (jj - 24) ** 2) / (2 * 4.0 ** 2))
    _, dobs = _oracle_misfit_force(vp.astype(np.float64) * (1 + 0.08 * bump), vs.astype(np.float64),
                                   rho.astype(np.float64) * (1 - 0.05 * bump), m, rec, p, None, shots, amps)
    data, _ = m.model(torch.from_numpy(vp), torch.from_numpy(vs), torch.from_numpy(rho))
    _, ref = _oracle_misfit_force(vp.astype(np.float64), vs.astype(np.float64), rho.astype(np.float64), m, rec,
                                  p, None, shots, amps)
    assert np.linalg.norm(data.cpu().numpy() - ref) <= 1e-5 * np.linalg.norm(ref)
    tv = [torch.from_numpy(a).to("cuda").requires_grad_(True) for a in (vp, vs, rho)]
    loss = 0.5 * torch.sum((seismograms(*tv, m) - torch.from_numpy(dobs.astype(np.float32)).cuda()) ** 2)
    loss.backward()
    grads = [t.grad.cpu().numpy().astype(np.float64) for t in tv]
    base = [vp.astype(np.float64), vs.astype(np.float64), rho.astype(np.float64)]

    def fd(k, ci, cj):
        h = 1e-4 * base[k][ci, cj]
        mp = [b.copy() for b in base]
        mm = [b.copy() for b in base]
        mp[k][ci, cj] += h
        mm[k][ci, cj] -= h
        jp, _ = _oracle_misfit_force(*mp, m, rec, p, dobs, shots, amps)
        jm, _ = _oracle_misfit_force(*mm, m, rec, p, dobs, shots, amps)
        return (jp - jm) / (2 * h)

    rng = np.random.default_rng(3)
    cand = [(a, b) for a in range(12, shape[0] - 12) for b in range(12, shape[1] - 12)
            if max(abs(a - 30), abs(b - 14)) >= 4]
    cells = [cand[t] for t in rng.choice(len(cand), 12, replace=False)]
    for k, name in enumerate(("vp", "vs", "rho")):
        f = np.array([fd(k, ci, cj) for ci, cj in cells])
        a = np.array([grads[k][ci, cj] for ci, cj in cells])
        assert np.linalg.norm(a - f) <= 1e-3 * np.linalg.norm(f), (name, np.linalg.norm(a - f) / np.linalg.norm(f))
    f_src = fd(2, 30, 14)
    assert abs(grads[2][30, 14] - f_src) <= 1e-2 * abs(f_src), (grads[2][30, 14], f_src)


@pytest.mark.parametrize("source", ["force_vz", "force_vx"])
def test_force_source_batch_matches_shot_by_shot(source):
    from seiswave_taichi_amd.adjoint import seismograms
    shots = ((30, 14), (18, 20), (44, 16))
    amps = 1e6 * ricker2(np.arange(160), 8e-4, 1.5 / 20.0, 20.0, 1.0)
    mk = lambda: _modeling(shots=shots, ck=30)  # noqa: E731
    mb, _, _ = mk()
    ms, _, _ = mk()
    for m in (mb, ms):
        m.src_kind = {"force_vz": 1, "force_vx": 4}[source]
        m.force = True
        m.amp = torch.from_numpy(amps.astype(np.float32)).cuda()
    ms.batch = False
    vs, vp, rho = layered_model((mb.nx, mb.nz), seed=15)
    db, _ = mb.model(torch.from_numpy(vp), torch.from_numpy(vs), torch.from_numpy(rho))
    ds, _ = ms.model(torch.from_numpy(vp), torch.from_numpy(vs), torch.from_numpy(rho))
    np.testing.assert_array_equal(db.cpu().numpy(), ds.cpu().numpy())
    target = (0.9 * db).detach()
    grads = []
    for m in (mb, ms):
        tv = [torch.from_numpy(a).to("cuda").requires_grad_(True) for a in (vp, vs, rho)]
        loss = 0.5 * torch.sum((seismograms(*tv, m) - target) ** 2)
        loss.backward()
        grads.append([t.grad.double().cpu().numpy() for t in tv])
    for a, b in zip(*grads):
        assert np.linalg.norm(a - b) <= 1e-5 * np.linalg.norm(b)
    assert np.abs(grads[0][2][30, 14]) > 0
