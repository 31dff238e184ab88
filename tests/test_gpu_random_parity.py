"""Seeded random configurations through the three 3-D classes against the
oracle, bit for bit: grid shape, accuracy order 1-4, anisotropic spacing,
PML thickness and faces, source and receiver cells, class.  On grids this
small the automatic launch plan is one launch with CPML on every axis; the
five-launch plan (the CPML instantiated per box) is forced for the thin-PML
configurations (sw_set_launch_plan)."""
import numpy as np
import pytest

from oracle import oracle as O
from seiswave_taichi_amd.coeffs import diff_coff, ricker, ricker2
from tests.cases import layered_model, pml_dict, unit_mt

pytestmark = pytest.mark.gpu


def _case(seed):
    rng = np.random.default_rng(1000 + seed)
    kind = ("tensor", "cpml", "spml")[seed % 3]
    order = int(rng.integers(1, 5))
    npml = int(rng.integers(2, 7))
    lo = 2 * order + 2 * npml + 12
    shape = tuple(int(v) for v in rng.integers(lo, lo + 30, 3))
    d = tuple(float(v) for v in rng.choice([5.0, 7.5, 10.0, 12.5], 3))
    surface = [bool(b) for b in rng.random(6) < 0.8]
    vs, vp, rho = layered_model(shape, seed=seed)
    c = float(np.sum(np.abs(diff_coff(order))))
    dt = 0.8 / (float(vp.max()) * c * np.sqrt(np.sum(1.0 / np.square(d))))
    steps = int(rng.integers(20, 41))
    f = 0.25 / (steps * dt / 4)   # a few periods inside the run
    margin = order + 1
    src = tuple(int(rng.integers(margin + 5 if kind != "tensor" else margin, n - margin - 6)) for n in shape)
    rec = np.stack([rng.integers(0, n, 6) for n in shape], 1)
    return kind, order, npml, shape, d, surface, vs, vp, rho, dt, steps, f, src, rec


@pytest.mark.parametrize("seed", range(60))
def test_random_configuration_matches_oracle(seed):
    kind, order, npml, shape, d, surface, vs, vp, rho, dt, steps, f, src, rec = _case(seed)
    p = pml_dict(float(vp.max()), npml, d, f, float(vs.max()))
    if kind == "tensor":
        from seiswave_taichi_amd.Elastic3D_SSG_Tensor import ElasticWAVE
        m = unit_mt(seed)
        mt3 = np.zeros((3, 3), np.float32)
        mt3[0, 0], mt3[1, 1], mt3[2, 2], mt3[0, 1], mt3[0, 2], mt3[1, 2] = m
        w = ElasticWAVE(vs, vp, rho, *d, dt, *src, rec[:, 0][None], rec[:, 1][None], rec[:, 2][None], mt3, 1.0,
                        steps, order, f)
        o = O.Oracle3D("tensor", vs, vp, rho, *d, dt, src, rec, steps, order, freq=f, mt=m)
        amps = ricker2(np.arange(steps), dt, 0.03, f, 1.0)
    else:
        cls = "Elastic3D_SSG_CPML" if kind == "cpml" else "Elastic3D_SSG_SPML"
        ElasticWAVE = __import__(f"seiswave_taichi_amd.{cls}", fromlist=["ElasticWAVE"]).ElasticWAVE
        src = tuple(s - 5 for s in src)  # the force sits at (isx+5, isy+5, isz+5)
        scale = 1000.0 if kind == "cpml" else 3.0
        args = (vs, vp, rho, *d, dt, *src, *rec[0], steps, order, f)
        w = ElasticWAVE(*args) if kind == "cpml" else ElasticWAVE(*args, scale)
        o = O.Oracle3D(kind, vs, vp, rho, *d, dt, src, rec[:1], steps, order, freq=f, src_scale=scale)
        amps = ricker(np.arange(steps), dt, f, scale)
    w.SetADEPML3D(surface, p)
    o.set_pml(surface, p)
    if seed % 2:
        w.run(steps)
    else:
        for it in range(steps):
            w._run(it, 1)
    o.run(steps, amps=amps)
    for name in o.FIELDS:
        np.testing.assert_array_equal(getattr(w, name).to_numpy(), o.fields[name], err_msg=f"{name} {kind}")
    np.testing.assert_array_equal(w.data.to_numpy().reshape(o.data.shape), o.data)
    assert max(np.abs(o.fields[n]).max() for n in o.FIELDS) > 0


@pytest.fixture
def boxes_plan():
    """The five-launch plan wherever the bulk box is not empty, for one test."""
    from seiswave_taichi_amd import _lib as L
    lib = L.load()
    prev = lib.sw_set_launch_plan(2)  # SW_PLAN_BOXES
    yield
    lib.sw_set_launch_plan(prev)


@pytest.mark.parametrize("seed", range(6))
def test_random_five_launch_configuration_matches_oracle(seed, boxes_plan):
    """Seeded configurations whose PML is thin against the grid, with the
    five-launch plan forced: the CPML and Tensor classes run each sweep as
    five launches (CPML on x only in the bulk, x and y in the y slabs, all
    axes in the z slabs), the split-field class as a shell launch over its
    four boxes (thin x strips, y slabs, z slabs) and an interior launch;
    orders 1-4, bit for bit against the oracle, split fields included."""
    from seiswave_taichi_amd import _lib as L
    rng = np.random.default_rng(2000 + seed)
    kind = ("spml", "cpml", "tensor")[seed % 3]
    order = int(rng.integers(1, 5)) if kind != "cpml" else int(rng.integers(2, 5))
    npml = 2
    shape = tuple(int(v) for v in rng.integers(62, 72, 3))
    d = tuple(float(v) for v in rng.choice([7.5, 10.0, 12.5], 3))
    vs, vp, rho = layered_model(shape, seed=50 + seed)
    c = float(np.sum(np.abs(diff_coff(order))))
    dt = 0.8 / (float(vp.max()) * c * np.sqrt(np.sum(1.0 / np.square(d))))
    steps = 30
    f = 0.25 / (steps * dt / 4)
    # a source near a corner, so the signal crosses into the slabs of all three axes
    src = tuple(int(order + 3 + rng.integers(0, 4)) for _ in shape)
    rec = np.stack([rng.integers(0, n, 6) for n in shape], 1)
    p = pml_dict(float(vp.max()), npml, d, f, float(vs.max()))
    if kind == "tensor":
        from seiswave_taichi_amd.Elastic3D_SSG_Tensor import ElasticWAVE
        m = unit_mt(seed)
        mt3 = np.zeros((3, 3), np.float32)
        mt3[0, 0], mt3[1, 1], mt3[2, 2], mt3[0, 1], mt3[0, 2], mt3[1, 2] = m
        w = ElasticWAVE(vs, vp, rho, *d, dt, *src, rec[:, 0][None], rec[:, 1][None], rec[:, 2][None], mt3, 1.0,
                        steps, order, f)
        o = O.Oracle3D("tensor", vs, vp, rho, *d, dt, src, rec, steps, order, freq=f, mt=m)
        amps = ricker2(np.arange(steps), dt, 0.03, f, 1.0)
    else:
        cls = "Elastic3D_SSG_CPML" if kind == "cpml" else "Elastic3D_SSG_SPML"
        ElasticWAVE = __import__(f"seiswave_taichi_amd.{cls}", fromlist=["ElasticWAVE"]).ElasticWAVE
        if kind == "cpml":
            src = tuple(s_ - 5 for s_ in src)  # the force sits at (isx+5, isy+5, isz+5)
        scale = 1000.0 if kind == "cpml" else 3.0
        args = (vs, vp, rho, *d, dt, *src, *rec[0], steps, order, f)
        w = ElasticWAVE(*args) if kind == "cpml" else ElasticWAVE(*args, scale)
        o = O.Oracle3D(kind, vs, vp, rho, *d, dt, src, rec[:1], steps, order, freq=f, src_scale=scale)
        amps = ricker(np.arange(steps), dt, f, scale)
    w.SetADEPML3D([True] * 6, p)
    o.set_pml([True] * 6, p)
    assert L.launch_plan(w._e.descriptor()) == ([7, 0] if kind == "spml" else [1, 3, 3, 7, 7])
    w.run(steps)
    o.run(steps, amps=amps)
    for name in o.FIELDS:
        np.testing.assert_array_equal(getattr(w, name).to_numpy(), o.fields[name], err_msg=f"{name} {kind}")
    if kind == "spml":
        from seiswave_taichi_amd._classes3d import SPLIT_NAMES
        for i, name in enumerate(SPLIT_NAMES):
            np.testing.assert_array_equal(w.split_field(name), o.split[i], err_msg=name)
        assert sum(np.count_nonzero(q) for q in o.split) > 0
    np.testing.assert_array_equal(w.data.to_numpy().reshape(o.data.shape), o.data)
    assert max(np.abs(o.fields[n]).max() for n in o.FIELDS) > 0
