"""Snapshots taken off the step stream equal the fields read synchronously at
the same steps (the reference drivers' ``to_numpy()`` every 20 steps)."""
import numpy as np
import pytest

from tests.cases import layered_model, pml_dict, unit_mt

pytestmark = pytest.mark.gpu


def _tensor(steps):
    from seiswave_taichi_amd.Elastic3D_SSG_Tensor import ElasticWAVE
    shape = (40, 36, 44)
    vs, vp, rho = layered_model(shape, seed=21)
    m = unit_mt(4)
    mt3 = np.zeros((3, 3), np.float32)
    mt3[0, 0], mt3[1, 1], mt3[2, 2], mt3[0, 1], mt3[0, 2], mt3[1, 2] = m
    rx = np.arange(5, 35, 5)
    w = ElasticWAVE(vs, vp, rho, 10.0, 10.0, 10.0, 5e-4, 20, 18, 22, rx[None], np.full_like(rx, 18)[None],
                    np.full_like(rx, 6)[None], mt3, 1.0, steps, 2, 20.0)
    w.SetADEPML3D([True] * 6, pml_dict(float(vp.max()), 5, (10.0,) * 3, 20.0, float(vs.max())))
    return w


@pytest.mark.parametrize("index", [None, (slice(None), slice(None), 22)])
def test_3d_snapshots_match_synchronous_reads(index):
    from seiswave_taichi_amd.snapshots import Snapshots
    steps, every = 75, 20
    a, b = _tensor(steps), _tensor(steps)
    frames = Snapshots(a, fields=("vz", "sxx"), every=every, index=index, slots=2).run(steps, 0)
    ref = []
    for it in range(steps):
        b.update_SSG(it)
        if it % every == 0:
            ref.append((it, {n: getattr(b, n).to_numpy() if index is None else getattr(b, n).to_numpy()[index]
                             for n in ("vz", "sxx")}))
    assert [k for k, _ in frames] == [0, 20, 40, 60]
    for (ka, fa), (kb, fb) in zip(frames, ref):
        assert ka == kb
        for n in fb:
            np.testing.assert_array_equal(fa[n], fb[n], err_msg=f"{n} at step {ka}")
    assert np.abs(frames[-1][1]["vz"]).max() > 0
    np.testing.assert_array_equal(a.vz.to_numpy(), b.vz.to_numpy())


def test_2d_snapshots_match_synchronous_reads():
    from seiswave_taichi_amd.Elastic2D_CPML import ElasticWAVE
    from seiswave_taichi_amd.snapshots import Snapshots
    shape, steps = (60, 50), 64
    vs, vp, rho = layered_model(shape, seed=22)
    mk = lambda: ElasticWAVE(vs, vp, rho, 10.0, 10.0, 6e-4, 30, 25, 20, 8, steps, 2, 25.0)  # noqa: E731
    a, b = mk(), mk()
    for w in (a, b):
        w.SetADEPML2D([True] * 4, pml_dict(float(vp.max()), 6, (10.0, 10.0), 25.0, float(vs.max()), dims=2))
    frames = Snapshots(a, fields=("vx",), every=16, scheme="VS").run(steps - 5, 5)
    for it in range(steps):
        if it < 5:
            continue
        b.update_SSG_VS(it)
        if it % 16 == 0:
            got = dict(frames)[it]["vx"]
            np.testing.assert_array_equal(got, b.vx.to_numpy(), err_msg=str(it))
    assert [k for k, _ in frames] == [16, 32, 48]


def test_spml_snapshots_match_oracle():
    """The snapshot cadence of the reference's split-field driver
    (test3d_homegenous_SPML.py:71-80: update_SSG_SPML(i), then every 20 steps
    vx.to_numpy() and its y = isy plane): the frames taken off the step
    stream equal the oracle's vx plane after the same steps, bit for bit,
    and so does the whole field at the end."""
    from oracle import oracle as O
    from seiswave_taichi_amd.Elastic3D_SSG_SPML import ElasticWAVE
    from seiswave_taichi_amd.coeffs import ricker
    from seiswave_taichi_amd.snapshots import Snapshots
    shape, steps, every, f, dt = (44, 40, 42), 101, 20, 25.0, 4e-4
    vs, vp, rho = layered_model(shape, seed=23)
    isx, isy, isz = 22, 20, 21
    rs = (12, 10, 11)
    w = ElasticWAVE(vs, vp, rho, 10.0, 10.0, 10.0, dt, isx, isy, isz, *rs, steps, 2, f, 3.0)
    p = pml_dict(float(vp.max()), 6, (10.0,) * 3, f, float(vs.max()))
    w.SetADEPML3D([True] * 6, p)
    frames = Snapshots(w, fields=("vx",), every=every, index=(slice(None), isy, slice(None))).run(steps, 0)
    o = O.Oracle3D("spml", vs, vp, rho, 10.0, 10.0, 10.0, dt, (isx, isy, isz), [rs], steps, 2, freq=f, src_scale=3.0)
    o.set_pml([True] * 6, p)
    amps = ricker(np.arange(steps), dt, f, 3.0)
    ref = []
    for it in range(steps):
        o.step(it, float(amps[it]))
        if it % every == 0:
            ref.append((it, o.fields["vx"][:, isy, :].copy()))
    assert [k for k, _ in frames] == [k for k, _ in ref] == [0, 20, 40, 60, 80, 100]
    for (ka, fa), (_, fb) in zip(frames, ref):
        np.testing.assert_array_equal(fa["vx"], fb, err_msg=f"vx plane at step {ka}")
    assert np.abs(ref[-1][1]).max() > 0
    np.testing.assert_array_equal(w.vx.to_numpy(), o.fields["vx"])
    np.testing.assert_array_equal(w.data.to_numpy(), o.data[:, 0])


def test_2d_snapshots_match_oracle():
    """2-D VS run with snapshots of vz and sxx every 16 steps: each frame
    equals the oracle's field after that step."""
    from oracle import oracle as O
    from seiswave_taichi_amd.Elastic2D_CPML import ElasticWAVE
    from seiswave_taichi_amd.coeffs import ricker2
    from seiswave_taichi_amd.snapshots import Snapshots
    shape, steps, f, dt = (60, 50), 70, 25.0, 6e-4
    vs, vp, rho = layered_model(shape, seed=24)
    w = ElasticWAVE(vs, vp, rho, 10.0, 10.0, dt, 30, 25, 20, 8, steps, 2, f)
    p = pml_dict(float(vp.max()), 6, (10.0, 10.0), f, float(vs.max()), dims=2)
    w.SetADEPML2D([True] * 4, p)
    frames = dict(Snapshots(w, fields=("vz", "sxx"), every=16, scheme="VS").run(steps, 0))
    o = O.Oracle2D(vs, vp, rho, 10.0, 10.0, dt, 30, 25, 20, 8, steps, 2, f)
    o.set_pml([True] * 4, p)
    amps = ricker2(np.arange(steps), dt, 1.5 / f, f, 1.0)
    seen = 0
    for it in range(steps):
        o.vs_step(it, float(amps[it]))
        if it % 16 == 0:
            for n in ("vz", "sxx"):
                np.testing.assert_array_equal(frames[it][n], o.fields[n], err_msg=f"{n} at step {it}")
            seen += 1
    assert seen == 5 and np.abs(frames[64]["vz"]).max() > 0
