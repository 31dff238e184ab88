"""Pins of the CPU oracle (no GPU needed).

The reference ships no tests, golden vectors or fixtures for this path, so the
oracle is pinned by exact rationals, closed forms, invariants, the literal
constants the reference source holds, and an independent NumPy restatement.
"""
import math

import numpy as np
import pytest

from oracle import oracle as O
from oracle.numpy_ref import NumpyRef2D, NumpyRef3D
from tests.cases import homogeneous, layered_model, pml_dict, rel_l2

# Taylor staggered coefficients as exact rationals (src/Elastic3D_SSG_CPML.py:311-322
# solves for them; standard result, e.g. Levander 1988 for N=2).
RATIONALS = {
    1: [1.0],
    2: [9 / 8, -1 / 24],
    3: [75 / 64, -25 / 384, 3 / 640],
    4: [1225 / 1024, -245 / 3072, 49 / 5120, -5 / 7168],
    5: [19845 / 16384, -735 / 8192, 567 / 40960, -405 / 229376, 35 / 294912],
}


@pytest.mark.parametrize("order", sorted(RATIONALS))
def test_diff_coff_exact_rationals(order):
    c = O.diff_coff(order)
    np.testing.assert_allclose(c, RATIONALS[order], rtol=1e-13, atol=0)


@pytest.mark.parametrize("order", range(1, 9))
def test_diff_coff_moment_conditions(order):
    c = O.diff_coff(order)
    j = np.arange(order)
    for i in range(order):
        m = np.sum(c * (2 * j + 1.0) ** (2 * i + 1))
        assert abs(m - (1.0 if i == 0 else 0.0)) < 1e-9 * max(1.0, (2 * order - 1.0) ** (2 * i + 1))


def _poly_case(order, deg, axis_field, axis, kind="cpml"):
    """f64 oracle, unit rho and dt, field = x**deg along ``axis``: one velocity
    sweep returns the staggered derivative at the half points."""
    n = 4 * order + 14
    vs, vp, rho = homogeneous((n, n, n), vp=1.0, vs=0.5, rho=1.0)
    d = (0.7, 1.3, 0.9)
    o = O.Oracle3D(kind, vs, vp, rho, d[0], d[1], d[2], 1.0, (n // 2,) * 3, [(1, 1, 1)], 1, order,
                   precision="f64")
    o.set_pml([False] * 6, {"vp_max": 1.0, "pml_x_thick": 1e-3, "pml_y_thick": 1e-3, "pml_z_thick": 1e-3,
                            "Rcoef": 1e-6, "theta": 1.0, "alpha_max_pml": 0.0, "kmax_pml": 1.0})
    idx = np.arange(n, dtype=np.float64) * d[axis]
    shp = [1, 1, 1]
    shp[axis] = n
    o.fields[axis_field][...] = (idx ** deg).reshape(shp)
    o.velocity()
    return o, d, n


@pytest.mark.parametrize("order", [1, 2, 3, 4])
def test_staggered_stencil_exact_on_polynomials(order):
    """Dx3fm/dx of x**m at x_{i+1/2} equals m x^(m-1) for m <= 2N (vx loop)."""
    for deg in range(0, 2 * order + 1):
        o, d, n = _poly_case(order, deg, "sxx", 0)
        s = order
        i = np.arange(s, n - s - 1)
        xh = (i + 0.5) * d[0]
        expect = deg * xh ** (deg - 1) if deg > 0 else np.zeros_like(xh)
        got = o.fields["vx"][s:n - s - 1, n // 2, n // 2]
        np.testing.assert_allclose(got, expect, rtol=1e-9, atol=1e-9 * max(1.0, np.abs(expect).max()))


def test_staggered_stencil_not_exact_beyond_order():
    """A degree-(2N+1) polynomial is not differentiated exactly (the test above
    would catch a wrong coefficient)."""
    order = 2
    o, d, n = _poly_case(order, 2 * order + 1, "sxx", 0)
    i = np.arange(order, n - order - 1)
    xh = (i + 0.5) * d[0]
    got = o.fields["vx"][order:n - order - 1, n // 2, n // 2]
    assert np.abs(got - 5 * xh ** 4).max() > 1e-6


@pytest.mark.parametrize("axis,field,comp", [(1, "syy", "vy"), (2, "szz", "vz")])
def test_forward_stencils_other_axes(axis, field, comp):
    order = 3
    o, d, n = _poly_case(order, 5, field, axis)
    s = order
    i = np.arange(s, n - s - 1)
    xh = (i + 0.5) * d[axis]
    sl = [n // 2] * 3
    sl[axis] = slice(s, n - s - 1)
    got = o.fields[comp][tuple(sl)]
    np.testing.assert_allclose(got, 5 * xh ** 4, rtol=1e-9)


def test_wavelets_closed_form():
    dt, f = 1e-3, 15.0
    for n in range(0, 200, 7):
        t = n * dt
        a = math.pi ** 2 * f ** 2
        tau = t - 1.5 / f
        r1 = -2 * a * tau * math.exp(-a * tau * tau)
        tau2 = t - 0.03
        r2 = (1 - 2 * a * tau2 ** 2) * math.exp(-a * tau2 ** 2)
        assert abs(O.ricker(n, dt, f, 1.0, "f64") - r1) <= 1e-9 * max(1, abs(r1))
        assert abs(O.ricker2(n, dt, 0.03, f, 1.0, "f64") - r2) <= 1e-12
        assert abs(O.ricker(n, dt, f, 1.0, "f32") - r1) <= 2e-5 * 2 * a * 0.1
        assert abs(O.ricker2(n, dt, 0.03, f, 1.0, "f32") - r2) <= 2e-6


def test_product_wavelets_match_oracle_f32():
    from seiswave_taichi_amd.coeffs import ricker, ricker2
    dt, f = 1e-3, 10.0
    n = np.arange(1500)
    p1 = ricker(n, dt, f, 1000.0)
    p2 = ricker2(n, dt, 0.03, f, 1.0)
    o1 = np.array([O.ricker(int(i), dt, f, 1000.0) for i in n], np.float32)
    o2 = np.array([O.ricker2(int(i), dt, 0.03, f, 1.0) for i in n], np.float32)
    # glibc expf is not correctly rounded in every case; the product rounds an f64 exp.
    for p, o in ((p1, o1), (p2, o2)):
        ulp = np.abs(p.view(np.int32).astype(np.int64) - o.view(np.int32).astype(np.int64))
        assert ulp.max() <= 1


PML_CASES = [
    dict(n=60, s=2, d=10.0, dt=1e-3, npml=20, vp=3500.0, f=10.0, vsmax=2000.0, lo=True, hi=True, theta=1.0),
    dict(n=47, s=3, d=12.5, dt=5.9e-4, npml=10, vp=5500.0, f=10.0, vsmax=3179.0, lo=True, hi=False, theta=0.5),
    dict(n=40, s=5, d=2.5, dt=3e-4, npml=15, vp=3000.0, f=60.0, vsmax=1796.0, lo=False, hi=True, theta=1.0),
    dict(n=33, s=2, d=1, dt=1e-4, npml=8, vp=2000.0, f=30.0, vsmax=1150.0, lo=True, hi=True, theta=0.0),
]


@pytest.mark.parametrize("case", PML_CASES)
def test_pml_profile_product_equals_oracle(case):
    from seiswave_taichi_amd.pml import axis_pml
    c = case
    thick = c["npml"] * c["d"]
    args = (c["n"], c["s"], c["d"], c["dt"], thick, c["vp"], 1e-6, math.pi * c["f"],
            c["vsmax"] / (5 * c["d"] * c["f"]), c["theta"], c["lo"], c["hi"])
    ref, mn, mx = O.pml_axis(*args)
    got = axis_pml(*args)
    assert got.min_pml == mn and got.max_pml == mx
    for k, v in ref.items():
        np.testing.assert_array_equal(getattr(got, k), v, err_msg=k)


@pytest.mark.parametrize("case", PML_CASES)
def test_pml_profile_closed_forms(case):
    c = case
    thick = c["npml"] * c["d"]
    kmax = c["vsmax"] / (5 * c["d"] * c["f"])
    amax = math.pi * c["f"]
    arrs, mn, mx = O.pml_axis(c["n"], c["s"], c["d"], c["dt"], thick, c["vp"], 1e-6, amax, kmax,
                              c["theta"], c["lo"], c["hi"], precision="f64")
    d0 = -3 * c["vp"] * math.log(1e-6) / (2 * thick)
    for half, sfx in ((0.0, ""), (c["d"] / 2, "_h")):
        x = (np.arange(c["n"]) - c["s"]) * c["d"] + half
        r = np.where((x <= mn) & c["lo"], (mn - x) / thick,
                     np.where((x >= mx) & c["hi"], (x - mx) / thick, np.nan))
        inside = ~np.isnan(r)
        rr = np.nan_to_num(r)
        np.testing.assert_allclose(arrs["d" + sfx], np.where(inside, d0 * rr ** 2, 0), rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(arrs["k" + sfx], np.where(inside, 1 + (kmax - 1) * rr ** 2, 1), rtol=1e-12)
        al = np.where(inside, np.maximum(amax * (1 - rr), 0), 0)
        np.testing.assert_allclose(arrs["alpha" + sfx], al, rtol=1e-12, atol=1e-12)
        D, K, A = arrs["d" + sfx], arrs["k" + sfx], arrs["alpha" + sfx]
        q = D / K + A
        th = c["theta"]
        np.testing.assert_allclose(arrs["b" + sfx], (1 - (1 - th) * c["dt"] * q) / (1 + th * c["dt"] * q),
                                   rtol=1e-12)
        aexp = np.where(np.abs(D) > 1e-6, -c["dt"] * D / K ** 2 / (1 + th * c["dt"] * q), 0)
        np.testing.assert_allclose(arrs["a" + sfx], aexp, rtol=1e-12, atol=1e-300)
        # outside the PML the coefficients are the identity
        assert np.all(arrs["b" + sfx][~inside] == 1) and np.all(arrs["a" + sfx][~inside] == 0)


def test_slab_bounds_twenty_cell_pml():
    """With N=2 and a 20-cell PML the non-identity coefficients occupy indices
    0..21 on the low side and n-22..n-1 on the high side."""
    from seiswave_taichi_amd.pml import axis_pml
    n = 200
    ax = axis_pml(n, 2, 10.0, 1e-3, 200.0, 3500.0, 1e-6, math.pi * 10, 4.0, 1.0, True, True)
    assert (ax.lo, ax.hi) == (22, n - 23)
    nontriv = (ax.a != 0) | (ax.k != 1) | (ax.a_h != 0) | (ax.k_h != 1)
    assert not nontriv[ax.lo:ax.hi].any() and nontriv[ax.lo - 1] and nontriv[ax.hi]


# ----------------------------------------------------- oracle vs NumPy, f64
def _ref3d(kind, shape, order, seed, steps, npml=4, surface=(True,) * 6, d=(10.0, 12.5, 8.0)):
    vs, vp, rho = layered_model(shape, seed=seed)
    dt = 5e-4
    src = (shape[0] // 2, shape[1] // 2 - 1, shape[2] // 2 + 1)
    rec = [(order + 3, shape[1] // 2, order + 2), (shape[0] - order - 3, 3 + order, shape[2] // 2)]
    mt = [0.3, -0.2, 0.5, 0.4, -0.1, 0.25]
    o = O.Oracle3D(kind, vs, vp, rho, *d, dt, src, rec, steps, order, freq=20.0, src_scale=1.0, mt=mt,
                   precision="f64")
    p = pml_dict(float(vp.max()), npml, d, 20.0, float(vs.max()))
    o.set_pml(list(surface), p)
    pml = [o.pml[a] for a in "xyz"]
    bounds = [(o.st.xmin_pml, o.st.xmax_pml), (o.st.ymin_pml, o.st.ymax_pml), (o.st.zmin_pml, o.st.zmax_pml)]
    nref = NumpyRef3D(kind, o.rho, o.mu, o.lam, *d, dt, o.c, pml, bounds,
                      o.src if kind == "tensor" else ((src[0] + 5, src[1] + 5, src[2] + 5) if kind == "cpml" else src),
                      mt=mt, rec=[tuple(r) for r in rec])
    # the CPML class injects at (isx+5, isy+5, isz+5): keep the cell inside
    for it in range(steps):
        amp = o.wavelet(it)
        o.step(it, amp)
        nref.step(amp)
    return o, nref


@pytest.mark.parametrize("kind", ["cpml", "tensor", "spml"])
def test_oracle_matches_numpy_restatement_3d(kind):
    shape = (22, 20, 24) if kind != "cpml" else (26, 24, 28)
    o, n = _ref3d(kind, shape, 2, seed=1, steps=25)
    for f in o.FIELDS:
        assert rel_l2(o.fields[f], n.f[f]) < 1e-12, f
    np.testing.assert_allclose(o.data, np.array(n.data), rtol=1e-10, atol=1e-14 * np.abs(o.data).max())


def test_oracle_matches_numpy_restatement_3d_order3_partial_pml():
    o, n = _ref3d("tensor", (24, 22, 26), 3, seed=2, steps=20, npml=5,
                  surface=(True, False, False, True, True, True))
    for f in o.FIELDS:
        assert rel_l2(o.fields[f], n.f[f]) < 1e-12, f


@pytest.mark.parametrize("scheme", ["VS", "SV", "RSG"])
def test_oracle_matches_numpy_restatement_2d(scheme):
    shape = (40, 36)
    vs, vp, rho = layered_model(shape, seed=3)
    d = (10.0, 7.5)
    dt = 6e-4
    order = 3
    o = O.Oracle2D(vs, vp, rho, d[0], d[1], dt, 20, 18, 9, 6, 40, order, freq=25.0, precision="f64")
    p = pml_dict(float(vp.max()), 6, d, 25.0, float(vs.max()), dims=2)
    o.set_pml([True, True, True, True], p)
    nref = NumpyRef2D(scheme, o.rho, o.mu, o.lam, d[0], d[1], dt, o.c, [o.pml["x"], o.pml["z"]],
                      [(o.st.xmin_pml, o.st.xmax_pml), (o.st.zmin_pml, o.st.zmax_pml)], (20, 18), (9, 6))
    for it in range(40):
        if scheme == "VS":
            amp = O.ricker2(it, dt, 1.5 / 25.0, 25.0, 1.0, "f64")
            o.vs_step(it, amp)
        elif scheme == "SV":
            amp = O.ricker(it, dt, 25.0, 1.0, "f64")
            o.sv_step(it, amp)
        else:
            amp = O.ricker(it, dt, 25.0, 1.0, "f64")
            o.rsg_step(amp)
        nref.step(amp)
    for f in o.FIELDS:
        assert rel_l2(o.fields[f], nref.f[f]) < 1e-12, f


# ---------------------------------------------------------- invariants
def test_causality_light_cone():
    """Each step moves information at most 2N cells (N per sweep): cells farther
    than 2N*steps+1 from the source stay exactly zero."""
    shape = (40, 40, 40)
    vs, vp, rho = homogeneous(shape)
    o = O.Oracle3D("tensor", vs, vp, rho, 10.0, 10.0, 10.0, 1e-3, (20, 20, 20), [(5, 5, 5)], 4, 2,
                   freq=10.0, mt=[1, 1, 1, 0, 0, 0], precision="f64")
    o.set_pml([True] * 6, pml_dict(3500.0, 5, (10.0,) * 3, 10.0, 2000.0))
    for it in range(4):
        o.step(it)
    reach = 2 * 2 * 4 + 1
    ii, jj, kk = np.meshgrid(*(np.arange(n) for n in shape), indexing="ij")
    far = np.maximum.reduce([np.abs(ii - 20), np.abs(jj - 20), np.abs(kk - 20)]) > reach
    for f in o.FIELDS:
        assert np.all(o.fields[f][far] == 0), f
        assert np.any(o.fields[f] != 0), f


def test_linearity_in_source_amplitude():
    shape = (24, 24, 24)
    vs, vp, rho = layered_model(shape, seed=4)
    outs = []
    for scale in (1.0, 3.0):
        o = O.Oracle3D("spml", vs, vp, rho, 10.0, 10.0, 10.0, 5e-4, (12, 12, 12), [(4, 12, 4)], 30, 2,
                       freq=20.0, src_scale=scale, precision="f64")
        o.set_pml([True] * 6, pml_dict(float(vp.max()), 5, (10.0,) * 3, 20.0, float(vs.max())))
        o.run(30)
        outs.append(o.fields["vz"].copy())
    np.testing.assert_allclose(outs[1], 3.0 * outs[0], rtol=1e-12, atol=1e-30)


def test_explosive_source_symmetry():
    """MT = I/sqrt(3) in a homogeneous medium: szz after n steps is symmetric
    under x <-> y about the source (the grid is cubic and the staggering of
    szz is the same along x and y)."""
    n = 30
    vs, vp, rho = homogeneous((n, n, n))
    m = 1 / math.sqrt(3)
    o = O.Oracle3D("tensor", vs, vp, rho, 10.0, 10.0, 10.0, 1e-3, (15, 15, 15), [(5, 5, 5)], 12, 2,
                   freq=15.0, mt=[m, m, m, 0, 0, 0], precision="f64")
    o.set_pml([True] * 6, pml_dict(3500.0, 5, (10.0,) * 3, 15.0, 2000.0))
    o.run(12)
    szz = o.fields["szz"]
    np.testing.assert_allclose(szz, szz.transpose(1, 0, 2), rtol=1e-12, atol=1e-12 * np.abs(szz).max())
    sxx, syy = o.fields["sxx"], o.fields["syy"]
    np.testing.assert_allclose(sxx, syy.transpose(1, 0, 2), rtol=1e-9, atol=1e-9 * np.abs(sxx).max())


def test_p_wave_arrival_time():
    """Explosive source in a homogeneous medium: the P wave reaches a receiver
    at distance r near t = t_peak + r/vp (within two periods of grid dispersion)."""
    n = 60
    vp0 = 3000.0
    vs, vp, rho = homogeneous((n, n, n), vp=vp0, vs=1732.0, rho=2000.0)
    dt, f = 1e-3, 15.0
    m = 1 / math.sqrt(3)
    src = (30, 30, 30)
    rec = [(30, 30 + 15, 30)]
    nt = 200
    o = O.Oracle3D("tensor", vs, vp, rho, 10.0, 10.0, 10.0, dt, src, rec, nt, 2, freq=f,
                   mt=[m, m, m, 0, 0, 0], precision="f32")
    o.set_pml([True] * 6, pml_dict(vp0, 10, (10.0,) * 3, f, 1732.0))
    o.run(nt)
    tr = np.abs(o.data[:, 0])
    t_peak_obs = np.argmax(tr) * dt
    t_expect = 0.03 + 150.0 / vp0
    assert abs(t_peak_obs - t_expect) < 1.5 / f


def test_pml_absorbs_energy():
    """With the 20-cell-style CPML, the wavefield energy decays once the wave
    reaches the boundary; with PML off (reflecting), it stays high."""
    n = 40
    vs, vp, rho = homogeneous((n, n, n))
    energies = {}
    for on in (True, False):
        o = O.Oracle3D("cpml", vs, vp, rho, 10.0, 10.0, 10.0, 1e-3, (15, 15, 15), [(5, 5, 5)], 300, 2,
                       freq=25.0, precision="f32")
        o.set_pml([on] * 6, pml_dict(3500.0, 8, (10.0,) * 3, 25.0, 2000.0))
        o.run(300)
        energies[on] = sum(float(np.sum(o.fields[f].astype(np.float64) ** 2)) for f in ("vx", "vy", "vz"))
    assert energies[True] < 0.05 * energies[False]


def test_s_wave_arrival_time():
    """2-D vertical force in a homogeneous medium: along the horizontal through
    the source the far field is pure S, so vz there peaks near
    t0 + r/vs (t0 = 1.5/f, the Ricker2 delay of update_SSG_VS)."""
    n, vp0, vs0, f, dt = 120, 3000.0, 1732.0, 15.0, 1e-3
    vs, vp, rho = homogeneous((n, n), vp=vp0, vs=vs0, rho=2000.0)
    nt, r = 330, 20
    o = O.Oracle2D(vs, vp, rho, 10.0, 10.0, dt, 50, 60, 50 + r, 60, nt, 2, freq=f, precision="f32")
    o.set_pml([True] * 4, pml_dict(vp0, 20, (10.0, 10.0), f, vs0, dims=2))
    for it in range(nt):
        o.vs_step(it)
    t_peak_obs = np.argmax(np.abs(o.data)) * dt
    t_expect = 1.5 / f + r * 10.0 / vs0
    assert abs(t_peak_obs - t_expect) < 1.0 / f, (t_peak_obs, t_expect)
    # and not the P time: r/vp is 0.05 s earlier
    assert abs(t_peak_obs - (1.5 / f + r * 10.0 / vp0)) > 0.5 * abs(t_peak_obs - t_expect) + 0.02


def test_reciprocity_of_force_responses():
    """Source-receiver reciprocity of the discrete 2-D VS step in a layered
    medium (f64, no PML, before any boundary reflection returns): the vz
    response at B to a vz force at A equals the response at A to the force at
    B, once each is weighted by rho[src]/rho_vz[src] (the force adds
    dt*amp/rho at the integer cell while the vz update uses the four-point
    average at the vz point)."""
    shape = (90, 80)
    vs, vp, rho = layered_model(shape, seed=9)
    vs, vp, rho = (a.astype(np.float64) for a in (vs, vp, rho))
    a, b = (30, 33), (52, 45)
    nt, dt = 120, 5e-4
    traces = {}
    for s, r in ((a, b), (b, a)):
        o = O.Oracle2D(vs, vp, rho, 10.0, 10.0, dt, *s, *r, nt, 2, freq=25.0, precision="f64")
        o.set_pml([False] * 4, pml_dict(float(vp.max()), 5, (10.0, 10.0), 25.0, float(vs.max()), dims=2))
        for it in range(nt):
            o.vs_step(it)
        w = rho[s] / (0.25 * (rho[s] + rho[s[0], s[1] + 1] + rho[s[0] + 1, s[1] + 1] + rho[s[0] + 1, s[1]]))
        traces[s] = o.data * w
    ta, tb = traces[a], traces[b]
    assert np.abs(ta).max() > 0
    assert np.linalg.norm(ta - tb) <= 1e-10 * np.linalg.norm(ta), np.linalg.norm(ta - tb) / np.linalg.norm(ta)
