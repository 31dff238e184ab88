"""Pins of the oracle's split-field PML, 2-D SV and rotated-grid schemes by
what the mathematics fixes, written as functions that raise AssertionError on
a violation, so the same checks run against the committed oracle
(tests/test_oracle_pins.py) and against deliberately mistaken variants of
it (the mutation tests there).

Reference anchors: split-field PML src/Elastic3D_SSG_SPML.py:182-391;
update_SSG_SV src/Elastic2D_CPML.py:202-274; update_RSG :276-353 with the
rotated stencils src/Differential2D.py:32-56.
"""
import math

import numpy as np

from oracle import oracle as O
from tests.cases import homogeneous, layered_model, pml_dict

SPLITS = ("vx_x", "vx_y", "vx_z", "vy_x", "vy_y", "vy_z", "vz_x", "vz_y", "vz_z", "sxx_x", "sxx_y", "sxx_z",
          "syy_x", "syy_y", "syy_z", "sxy_x", "sxy_y", "sxz_x", "sxz_z", "syz_y", "syz_z", "szz_x", "szz_y",
          "szz_z")
# staggering of each field: True where the field sits at a half point of (x, y, z)
HALF3 = {"vx": (1, 0, 0), "vy": (0, 1, 0), "vz": (0, 0, 1), "sxx": (0, 0, 0), "syy": (0, 0, 0),
         "szz": (0, 0, 0), "sxy": (1, 1, 0), "sxz": (1, 0, 1), "syz": (0, 1, 1)}


def _mirror(a, axis, center2, half):
    """a reflected about the coordinate center2/2 (in cells) along axis: an
    integer point i maps to center2 - i, a half point i (at i + 1/2) to
    center2 - 1 - i.  Indices outside the array are dropped on both sides."""
    n = a.shape[axis]
    src = np.arange(n)
    dst = (center2 - 1 - src) if half else (center2 - src)
    ok = (dst >= 0) & (dst < n)
    sl_src = [slice(None)] * a.ndim
    sl_dst = [slice(None)] * a.ndim
    sl_src[axis] = src[ok]
    sl_dst[axis] = dst[ok]
    return a[tuple(sl_src)], a[tuple(sl_dst)]


# ----------------------------------------------------------- split-field PML
def _spml(shape, npml, nt, rcoef=1e-6, surface=(True,) * 6, src=None, d=10.0, dt=1e-3, f=20.0):
    vs, vp, rho = homogeneous(shape)
    src = src or tuple(n // 2 for n in shape)
    o = O.Oracle3D("spml", vs, vp, rho, d, d, d, dt, src, [(3, 3, 3)], nt, 2, freq=f, src_scale=1e6,
                   precision="f64")
    p = pml_dict(3500.0, npml, (d, d, d), f, 2000.0)
    p["Rcoef"] = rcoef
    o.set_pml(list(surface), p)
    return o


def spml_mirror_symmetry():
    """A vz force at the centre of a homogeneous box with the split-field PML
    on all faces: the solution is mirror symmetric about the source in x, y
    (vz even) and z (a z force: vz even, vx and vy odd), with every field of
    the parity its staggering and tensor character give it.  Each mirrored
    cell does the mirrored arithmetic (a - b becomes -(b - a), exact), so the
    symmetry holds bit for bit, also inside the PML where the damping of each
    split is looked up at the mirrored point (a d for d_h mistake breaks it).
    The update ranges are not mirror symmetric at the grid edges, so the run
    stops before the stencil light cone (2N cells per step) reaches them."""
    nxy, nz, npml, nt = 61, 60, 12, 9   # 9 steps: the fringe is 26 cells out, the PML 16
    o = _spml((nxy, nxy, nz), npml, nt, src=(30, 30, nz // 2 - 1))
    for it in range(nt):
        o.step(it, 1.0 if it < 3 else 0.0)
    # parity under (x, y, z) mirrors for a z force: x-odd / y-odd / z-odd
    odd = {"vx": (1, 0, 1), "vy": (0, 1, 1), "vz": (0, 0, 0), "sxx": (0, 0, 1), "syy": (0, 0, 1),
           "szz": (0, 0, 1), "sxy": (1, 1, 1), "sxz": (1, 0, 0), "syz": (0, 1, 0)}
    centers = (60, 60, nz - 1)  # twice the mirror coordinates: x = 30, y = 30, z = nz/2 - 1/2
    fields = dict(o.fields)
    fields.update({n: o.split[i] for i, n in enumerate(SPLITS)})
    inside = 0
    for name, a in fields.items():
        base = name.split("_")[0]
        for ax in range(3):
            x, y = _mirror(a, ax, centers[ax], HALF3[base][ax])
            sign = -1.0 if odd[base][ax] else 1.0
            assert np.array_equal(x, sign * y), f"{name} not mirror symmetric along axis {ax}"
    # the check reached the split-field region: damped split values are nonzero
    for i in range(24):
        inside += int(np.count_nonzero(o.split[i]))
    assert inside > 1000, inside


def spml_without_damping_is_the_plain_scheme():
    """With Rcoef = 1 the damping d = -3 vp ln(R)/(2L) r^2 is 0 everywhere, the
    Crank-Nicolson split update reduces to q += dt*(its derivative term) and
    the split sum to the plain update: the split-field class must equal the
    CPML class with every face off (identity CPML) up to f64 rounding, in a
    run long enough for the wave to cross the split region and return."""
    shape, npml, nt = (34, 30, 32), 6, 90
    o = _spml(shape, npml, nt, rcoef=1.0, src=(17, 15, 16))
    vs, vp, rho = homogeneous(shape)
    c = O.Oracle3D("cpml", vs, vp, rho, 10.0, 10.0, 10.0, 1e-3, (12, 10, 11), [(3, 3, 3)], nt, 2, freq=20.0,
                   precision="f64")
    c.set_pml([False] * 6, pml_dict(3500.0, npml, (10.0,) * 3, 20.0, 2000.0))
    for it in range(nt):
        amp = O.ricker(it, 1e-3, 20.0, 1e6, "f64")
        o.step(it, amp)
        c.step(it, amp)  # the CPML class injects at (isx+5, isy+5, isz+5) = (17, 15, 16)
    assert all(np.any(q != 0) for q in o.split[:24]), "some split component never updated"
    for name in o.FIELDS:
        a, b = o.fields[name], c.fields[name]
        assert np.abs(b).max() > 0
        err = np.linalg.norm(a - b) / np.linalg.norm(b)
        assert err < 1e-12, (name, err)


def spml_crank_nicolson_factors():
    """One velocity loop and one stress loop from a state with zero spatial
    derivatives and nonzero split values: each split q of a field obeys
    dq/dt = -d q there, whose Crank-Nicolson step multiplies q by
    (1 - dt d/2)/(1 + dt d/2), with d the damping profile of the split's own
    direction at the field's own (integer or half) point.  Every split of
    every field is checked against that factor at every cell it updates."""
    shape, npml = (26, 24, 22), 6
    o = _spml(shape, npml, 1)
    rng = np.random.default_rng(5)
    for q in o.split:
        q[...] = 1.0 + rng.random(q.shape)
    before = [q.copy() for q in o.split]
    o.velocity()   # stresses are 0: no derivative forcing of the velocity splits
    dt = 1e-3
    prof = {ax: (o.pml[ax]["d"], o.pml[ax]["d_h"]) for ax in "xyz"}
    s = 2

    def expect(base, direction, before_q):
        half = HALF3[base]["xyz".index(direction)]
        dd = prof[direction][1 if half else 0]
        shp = [1, 1, 1]
        shp["xyz".index(direction)] = len(dd)
        g = (1 - 0.5 * dt * dd) / (1 + 0.5 * dt * dd)
        return before_q * g.reshape(shp)

    # loop ranges of each field's update in the split-field class
    # (src/Elastic3D_SSG_SPML.py:182-391), as [lo, n - hi) per axis
    ranges = {"vx": ((s, s + 1), (s + 1, s), (s + 1, s)), "vy": ((s + 1, s), (s, s + 1), (s + 1, s)),
              "vz": ((s + 1, s), (s + 1, s), (s, s + 1)), "sxx": ((s + 1, s),) * 3, "syy": ((s + 1, s),) * 3,
              "szz": ((s + 1, s),) * 3, "sxy": ((s + 1, s), (s + 1, s), (s, s)),
              "sxz": ((s, s + 1), (s, s), (s, s + 1)), "syz": ((s, s), (s, s + 1), (s, s + 1))}

    def split_region(base):
        """Cells the reference's loop for `base` updates whose staggered point
        is in the PML (the coordinate test of src/Elastic3D_SSG_SPML.py:191)."""
        half = HALF3[base]
        m = np.zeros(shape, bool)
        (a0, a1), (b0, b1), (c0, c1) = ranges[base]
        m[a0:shape[0] - a1, b0:shape[1] - b1, c0:shape[2] - c1] = True
        inside = np.zeros(shape, bool)
        for ax, name in enumerate("xyz"):
            coord = (np.arange(shape[ax]) - s) * 10.0 + (5.0 if half[ax] else 0.0)
            mn, mx = getattr(o.st, f"{name}min_pml"), getattr(o.st, f"{name}max_pml")
            sel = (coord <= mn) | (coord >= mx)
            shp = [1, 1, 1]
            shp[ax] = shape[ax]
            inside |= sel.reshape(shp)
        return m & inside

    for i, name in enumerate(SPLITS[:9]):
        base, direction = name.split("_")
        reg = split_region(base)
        want = expect(base, direction, before[i])
        assert reg.sum() > 0
        np.testing.assert_allclose(o.split[i][reg], want[reg], rtol=1e-15, atol=0, err_msg=name)
    # stress splits: velocities zero (their splits zero), stresses split values nonzero
    o2 = _spml(shape, npml, 1)
    for q in o2.split[9:]:
        q[...] = 1.0 + rng.random(q.shape)
    before = [q.copy() for q in o2.split]
    o2.stress()
    for i, name in enumerate(SPLITS[9:], start=9):
        base, direction = name.split("_")
        reg = split_region(base)
        want = expect(base, direction, before[i])
        assert reg.sum() > 0
        np.testing.assert_allclose(o2.split[i][reg], want[reg], rtol=1e-15, atol=0, err_msg=name)


def spml_plane_wave_in_a_transverse_layer():
    """A plane wave travelling along x through a box whose only PML is on the
    y faces: where the fields do not depend on y (cells the y-edge effects
    have not reached), the y splits receive exactly zero, the x and z splits
    are undamped, and the split-field run equals the same run with no PML.
    A derivative routed into the split of another direction would be damped
    by that direction's profile and break this."""
    shape, npml, nt = (60, 56, 40), 20, 3
    runs = {}
    for surf in ((False, False, True, True, False, False), (False,) * 6):
        o = _spml(shape, npml, nt, surface=surf)
        x = np.arange(shape[0])[:, None, None]
        pulse = np.exp(-((x - 30.0) / 4.0) ** 2) * np.ones(shape)
        for name, amp in (("vx", 1.0), ("sxx", 2.0e6), ("syy", 0.5e6), ("szz", 0.5e6)):
            o.fields[name][...] = amp * pulse
        # the reference keeps a field as its split sum inside the PML: put the
        # initial values in the x splits
        for name, i in (("vx", 0), ("sxx", 9), ("syy", 12), ("szz", 21)):
            o.split[i][...] = o.fields[name]
        for it in range(nt):
            o.step(it, 0.0)
        runs[surf[2]] = o
    a, b = runs[True], runs[False]
    reach = 2 * 2 * nt + 2
    y0, y1 = 2 + 1 + reach, shape[1] - 2 - 1 - reach
    z0, z1 = 2 + 1 + reach, shape[2] - 2 - 1 - reach
    sel = (slice(6, 54), slice(y0, y1), slice(z0, z1))
    # the region reaches into the y PML (which starts 20 cells from each y face)
    assert y0 < 2 + npml and y1 > shape[1] - 2 - npml
    for name in a.FIELDS:
        ref = b.fields[name][sel]
        if np.abs(ref).max() == 0:
            continue
        err = np.abs(a.fields[name][sel] - ref).max() / np.abs(ref).max()
        assert err < 1e-13, (name, err)


def spml_moment_source_equals_tensor_before_the_boundary():
    """The moment-tensor source in the split-field class (an extension: the
    reference's split-field class has only a vz force) with M = I/sqrt(3) in
    a homogeneous medium: before the wave reaches the PML the split-field run
    equals the Tensor (CPML) class run, whose interior arithmetic is the same
    (point and averaged moduli coincide), up to f64 rounding and the
    numerical precursor of the stencil light cone."""
    n, npml, nt = 48, 8, 28
    vs, vp, rho = homogeneous((n, n, n))
    m = 1 / math.sqrt(3)
    mt = [m, m, m, 0.0, 0.0, 0.0]
    a = O.Oracle3D("spml", vs, vp, rho, 10.0, 10.0, 10.0, 1e-3, (24, 24, 24), [(24, 24, 14)], nt, 2, freq=15.0,
                   mt=mt, precision="f64")
    b = O.Oracle3D("tensor", vs, vp, rho, 10.0, 10.0, 10.0, 1e-3, (24, 24, 24), [(24, 24, 14)], nt, 2, freq=15.0,
                   mt=mt, precision="f64")
    p = pml_dict(3500.0, npml, (10.0,) * 3, 15.0, 2000.0)
    a.set_pml([True] * 6, p)
    b.set_pml([True] * 6, p)
    for it in range(nt):
        amp = O.ricker2(it, 1e-3, 1.5 / 15.0, 15.0, 1.0, "f64")
        a.step_moment(it, amp)
        b.step(it, amp)
    # the two classes differ only in the PML, which the light-cone precursor
    # (about 1e-9 of the field here) has entered
    assert np.abs(b.data).max() > 0
    for name in a.FIELDS:
        err = np.linalg.norm(a.fields[name] - b.fields[name]) / np.linalg.norm(b.fields[name])
        assert err < 1e-8, (name, err)
    assert np.linalg.norm(a.data - b.data) < 1e-8 * np.linalg.norm(b.data)


def spml_absorbs_energy():
    """Split-field PML on: the kinetic energy falls far below that of the same
    run in a rigid box (PML off) once the wave has reached the boundary."""
    n, nt = 40, 300
    energies = {}
    for on in (True, False):
        o = _spml((n, n, n), 8, nt, surface=(on,) * 6, src=(15, 15, 15), f=25.0)
        for it in range(nt):
            o.step(it)
        energies[on] = sum(float(np.sum(o.fields[f] ** 2)) for f in ("vx", "vy", "vz"))
    assert energies[True] < 0.05 * energies[False], energies


# ------------------------------------------------------------------- 2-D SV
def _sv(shape, src, rec, nt, npml=20, surface=(True,) * 4, model=None, f=15.0, dt=1e-3, d=10.0):
    vs, vp, rho = model if model is not None else homogeneous(shape, vp=3000.0, vs=1732.0, rho=2000.0)
    o = O.Oracle2D(vs, vp, rho, d, d, dt, *src, *rec, nt, 2, freq=f, precision="f64")
    o.set_pml(list(surface), pml_dict(float(np.max(vp)), npml, (d, d), f, float(np.max(vs)), dims=2))
    return o


def sv_s_wave_arrival():
    """update_SSG_SV's horizontal (vx) force in a homogeneous medium: on the
    vertical through the source the far field is pure S, so vx there peaks
    near t0 + r/vs (t0 = 1.5/f, the delay of the Ricker wavelet of
    src/BaseFun.py:4-11), within a period of the wavelet's own offset and grid
    dispersion, and well away from the P time r/vp."""
    n, f, r = 120, 15.0, 20
    o = _sv((n, n), (60, 40), (60, 40 + r), 330, f=f)
    for it in range(330):
        o.sv_step(it)
    t_peak = np.argmax(np.abs(o.data)) * 1e-3
    t_s = 1.5 / f + r * 10.0 / 1732.0
    t_p = 1.5 / f + r * 10.0 / 3000.0
    assert abs(t_peak - t_s) < 1.0 / f, (t_peak, t_s)
    assert abs(t_peak - t_p) > abs(t_peak - t_s) + 0.02, (t_peak, t_s, t_p)


def sv_reciprocity():
    """Source-receiver reciprocity of the discrete SV step (f64, layered
    medium, no PML, before any edge reflection returns): the vx response at B
    to a vx force at A equals the response at A to the force at B, each
    weighted by rho[src] / rho_x[src] (the force adds dt*amp/rho at the cell
    while the vx update divides by the x-averaged rho of
    src/Elastic2D_CPML.py:224)."""
    shape = (90, 80)
    vs, vp, rho = (a.astype(np.float64) for a in layered_model(shape, seed=9))
    a, b = (30, 33), (52, 45)
    traces = {}
    for s_, r_ in ((a, b), (b, a)):
        o = _sv(shape, s_, r_, 120, npml=5, surface=(False,) * 4, model=(vs, vp, rho), f=25.0, dt=5e-4)
        for it in range(120):
            o.sv_step(it)
        w = rho[s_] / (0.5 * (rho[s_] + rho[s_[0] + 1, s_[1]]))
        traces[s_] = o.data * w
    ta, tb = traces[a], traces[b]
    assert np.abs(ta).max() > 0
    assert np.linalg.norm(ta - tb) <= 1e-10 * np.linalg.norm(ta), np.linalg.norm(ta - tb) / np.linalg.norm(ta)


def sv_mirror_symmetry():
    """The SV force (vx at the half point (isx+1/2, isz)) in a homogeneous
    medium with CPML on all sides: mirror symmetric about x = isx + 1/2 (vx
    even, vz odd, sxx and szz even, sxz odd) and about z = isz (vx even, vz
    odd, sxx and szz even, sxz odd), bit for bit, through the CPML (each
    mirrored cell uses the mirrored coefficient: an integer/half mix-up breaks
    it), until the stencil light cone reaches the asymmetric edges."""
    nx, nz, nt = 80, 81, 12   # 12 steps: the fringe is 34 cells out, the CPML 25, the edges 37
    o = _sv((nx, nz), (39, 40), (3, 3), nt, npml=12)
    for it in range(nt):
        o.sv_step(it, 1.0 if it < 3 else 0.0)
    half = {"vx": (1, 0), "vz": (0, 1), "sxx": (0, 0), "szz": (0, 0), "sxz": (1, 1)}
    # an x force: x mirror keeps vx, flips vz, sxx, szz; z mirror flips vz, sxz
    odd = {"vx": (0, 0), "vz": (1, 1), "sxx": (1, 0), "szz": (1, 0), "sxz": (0, 1)}
    centers = (2 * 39 + 1, 2 * 40)
    for name in o.FIELDS:
        a = o.fields[name]
        for ax in range(2):
            x, y = _mirror(a, ax, centers[ax], half[name][ax])
            sign = -1.0 if odd[name][ax] else 1.0
            assert np.array_equal(x, sign * y), f"{name} not mirror symmetric along axis {ax}"
    nonzero_mem = sum(int(np.count_nonzero(m)) for m in o.mem)
    assert nonzero_mem > 500, nonzero_mem


# --------------------------------------------------------- 2-D rotated grid
def _rsg(n, npml, nt, f=20.0, dt=5e-4, d=10.0, vp=3000.0, vs=1732.0):
    vs_, vp_, rho = homogeneous((n, n), vp=vp, vs=vs, rho=2000.0)
    o = O.Oracle2D(vs_, vp_, rho, d, d, dt, n // 2, n // 2, 3, 3, nt, 2, freq=f, precision="f64")
    o.set_pml([True] * 4, pml_dict(vp, npml, (d, d), f, vs, dims=2))
    return o


def _explosive_start(o, c, width=2.5, amp=1e6):
    """sxx = szz = a Gaussian centred on the stress cell (c, c) (at c + 1/2)."""
    n = o.shape[0]
    x = np.arange(n) - c
    g = amp * np.exp(-(x[:, None] ** 2 + x[None, :] ** 2) / width ** 2)
    o.fields["sxx"][...] = g
    o.fields["szz"][...] = g


def rsg_transpose_symmetry():
    """Rotated grid (dx = dz, square grid, equal CPML on all sides) from an
    explosive start centred on a stress cell: swapping x and z maps the
    scheme onto itself.  The reference forms d/dx as (Drz - Drx)/(2 dx)
    (src/Elastic2D_CPML.py:299-302), the negative of the derivative along x,
    which is the physical system with vx and sxz negated, so the swap maps
    sxx to szz, sxz to itself and vx to -vz: bit for bit at every cell, at
    every step, CPML included (no light-cone limit: the update ranges are
    symmetric under the swap)."""
    n, nt = 64, 120
    o = _rsg(n, 10, nt)
    _explosive_start(o, 31)
    for it in range(nt):
        o.rsg_step(0.0)
        f = o.fields
        assert np.array_equal(f["sxx"], f["szz"].T), it
        assert np.array_equal(f["sxz"], f["sxz"].T), it
        assert np.array_equal(f["vx"], -f["vz"].T), it
    assert np.abs(o.fields["vx"]).max() > 0 and any(np.count_nonzero(m) for m in o.mem)


def rsg_p_arrival_along_a_diagonal():
    """Explosive start in a homogeneous medium: the pressure pulse reaches a
    point at distance r along the grid diagonal (where the rotated stencils
    act) near r/vp; the rotated derivative's 1/(2 dx) scale sets that speed
    (a wrong scale moves the arrival by a factor sqrt(2))."""
    n, nt, vp, dt = 140, 420, 3000.0, 5e-4
    o = _rsg(n, 12, nt, vp=vp, dt=dt)
    c = 69
    _explosive_start(o, c, width=2.0)
    k = 30                          # 30 cells along the diagonal: r = 30 sqrt(2) dx
    trace = []
    for it in range(nt):
        o.rsg_step(0.0)
        trace.append(o.fields["sxx"][c + k, c + k] + o.fields["szz"][c + k, c + k])
    t_peak = (np.argmax(np.abs(trace)) + 1) * dt
    t_p = k * math.sqrt(2) * 10.0 / vp
    assert abs(t_peak - t_p) < 0.1 * t_p, (t_peak, t_p)
    # and not at the S speed, nor at sqrt(2) times faster or slower
    assert abs(t_peak - t_p) < 0.5 * abs(t_peak - t_p * math.sqrt(2)), (t_peak, t_p)


def rsg_force_mirror_symmetry():
    """The reference's RSG force (vz += dt*source/(8 rho) on the four nodes
    around (isx - 1/2, isz - 1/2), src/Elastic2D_CPML.py:322-323) in a
    homogeneous medium with CPML: mirror symmetric about x = isx - 1/2 (vz
    even, vx odd, sxx and szz even, sxz odd), bit for bit, until the light
    cone reaches the asymmetric edges."""
    n, nt = 80, 12   # 12 steps: the fringe reaches 34 cells, the CPML 26, the edges 37
    vs_, vp_, rho = homogeneous((n, n), vp=3000.0, vs=1732.0, rho=2000.0)
    isx = isz = 40
    o = O.Oracle2D(vs_, vp_, rho, 10.0, 10.0, 5e-4, isx, isz, 3, 3, nt, 2, freq=20.0, precision="f64")
    o.set_pml([True] * 4, pml_dict(3000.0, 12, (10.0, 10.0), 20.0, 1732.0, dims=2))
    for it in range(nt):
        o.rsg_step(1e6 if it < 3 else 0.0)
    # velocities at integer nodes i, stresses at cells a (at a + 1/2):
    # node i <-> 2 isx - 1 - i, cell a <-> 2 isx - 2 - a
    half = {"vx": 0, "vz": 0, "sxx": 1, "szz": 1, "sxz": 1}
    odd = {"vx": 1, "vz": 0, "sxx": 0, "szz": 0, "sxz": 1}
    for name in o.FIELDS:
        x, y = _mirror(o.fields[name], 0, 2 * isx - 1, half[name])
        sign = -1.0 if odd[name] else 1.0
        assert np.array_equal(x, sign * y), f"{name} not mirror symmetric in x"
    assert sum(int(np.count_nonzero(m)) for m in o.mem) > 200
