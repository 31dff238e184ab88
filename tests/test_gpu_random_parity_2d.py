"""Seeded random configurations of the 2-D class against the oracle, bit for
bit: scheme (VS, SV, rotated grid), accuracy order 1-4, grid shape,
anisotropic spacing, PML thickness and faces, source kind (the reference's
force, explosive, moment tensor) and cell, receivers, and per-step calls
against multi-step native runs split at a random step."""
import numpy as np
import pytest

from oracle import oracle as O
from seiswave_taichi_amd.coeffs import diff_coff, ricker, ricker2
from tests.cases import layered_model, pml_dict

pytestmark = pytest.mark.gpu


def _case(seed):
    rng = np.random.default_rng(2000 + seed)
    scheme = ("VS", "SV", "RSG")[seed % 3]
    order = int(rng.integers(1, 5))
    npml = int(rng.integers(2, 9))
    lo = 2 * order + 2 * npml + 14
    shape = tuple(int(v) for v in rng.integers(lo, lo + 40, 2))
    d = tuple(float(v) for v in rng.choice([5.0, 7.5, 10.0, 12.5], 2))
    surface = [bool(b) for b in rng.random(4) < 0.8]
    vs, vp, rho = layered_model(shape, seed=seed)
    c = float(np.sum(np.abs(diff_coff(order))))
    dt = 0.8 / (float(vp.max()) * c * np.sqrt(np.sum(1.0 / np.square(d))))
    steps = int(rng.integers(30, 61))
    f = 0.25 / (steps * dt / 4)
    m = order + 2
    cell = tuple(int(rng.integers(m, n - m)) for n in shape)
    kind = ("force_vz", "explosive", "moment")[int(rng.integers(0, 3))] if scheme == "VS" else None
    rec = np.stack([rng.integers(0, n, 7) for n in shape], 1)
    split = int(rng.integers(1, steps)) if seed % 2 else None
    return scheme, order, npml, shape, d, surface, vs, vp, rho, dt, steps, f, cell, kind, rec, split


@pytest.mark.parametrize("seed", range(40))
def test_random_2d_configuration_matches_oracle(seed):
    from seiswave_taichi_amd.Elastic2D_CPML import ElasticWAVE
    scheme, order, npml, shape, d, surface, vs, vp, rho, dt, steps, f, cell, kind, rec, split = _case(seed)
    p = pml_dict(float(vp.max()), npml, d, f, float(vs.max()), dims=2)
    n = np.arange(steps)
    amps = ricker(n, dt, f, 1.0) if scheme != "VS" else ricker2(n, dt, 1.5 / f, f, 1.0)
    mt = np.array([0.6, 0.0, -0.3, 0.0, 0.45, 0.0], np.float32)
    source = None if kind is None else {"kind": kind, "cell": cell, "wavelet": amps, "mt": mt}
    kw = {} if scheme == "RSG" else {"receivers": rec, "components": ("vx", "vz"), "source": source}
    w = ElasticWAVE(vs, vp, rho, *d, dt, *cell, 0, 0, steps, order, f, **kw)
    w.SetADEPML2D(surface, p)
    o = O.Oracle2D(vs, vp, rho, *d, dt, *cell, 0, 0, steps, order, f)
    o.set_pml(surface, p)
    if scheme == "RSG":
        for it in range(steps):
            w.update_RSG(it, float(amps[it]))
    elif split is None:
        for it in range(steps):
            (w.update_SSG_VS if scheme == "VS" else w.update_SSG_SV)(it)
    else:
        w.run(scheme, split, 0)
        w.run(scheme, steps - split, split)
    ref = np.zeros((steps, 2, len(rec)), np.float32)
    for it in range(steps):
        a = np.float32(amps[it])
        if scheme == "RSG":
            o.rsg_step(float(a))
            continue
        if kind == "explosive":
            e = a * np.float32(-0.70710678118654752)
            o.fields["sxx"][cell] += e
            o.fields["szz"][cell] += e
        elif kind == "moment":
            o.fields["sxx"][cell] += a * -mt[0]
            o.fields["szz"][cell] += a * -mt[2]
            o.fields["sxz"][cell] += a * -mt[4]
        if scheme == "SV":
            o.sv_step(it, float(a))
        else:
            o.vs_step(it, float(a) if kind in (None, "force_vz") else 0.0)
        ref[it, 0] = o.fields["vx"][rec[:, 0], rec[:, 1]]
        ref[it, 1] = o.fields["vz"][rec[:, 0], rec[:, 1]]
    for name in o.FIELDS:
        np.testing.assert_array_equal(getattr(w, name).to_numpy(), o.fields[name], err_msg=f"{name} {scheme}")
    assert max(np.abs(o.fields[nm]).max() for nm in o.FIELDS) > 0
    if scheme != "RSG":
        np.testing.assert_array_equal(w.data.to_numpy().reshape(ref.shape), ref)
