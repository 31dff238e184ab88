"""The oracle's split-field PML, 2-D SV and rotated-grid restatements against
what the mathematics fixes (tests/oracle_pins.py), and proof that those pins
bite: each mutation test plants one plausible mistake in a copy of
oracle/seiswave_oracle.c (built into a temporary directory; the committed
oracle is untouched) and requires the named pins to reject it."""
import os
import shutil
import subprocess

import pytest

from oracle import oracle as O
from tests import oracle_pins as P

PINS = [n for n in dir(P) if n.startswith(("spml_", "sv_", "rsg_")) and callable(getattr(P, n))]
ORACLE_DIR = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.mark.parametrize("name", PINS)
def test_pin_holds_for_the_oracle(name):
    getattr(P, name)()


# (mistake, [(text in seiswave_oracle.c, replacement)], pins that must reject it)
MUTANTS = {
    "spml: vx x split damped with d at the integer point instead of d_h": (
        [("real pxn = (real)1 - hdt * st->px.d_h[i], pyn = (real)1 - hdt * st->py.d[j], pzn = (real)1 - hdt * st->pz.d[k];",
          "real pxn = (real)1 - hdt * st->px.d[i], pyn = (real)1 - hdt * st->py.d[j], pzn = (real)1 - hdt * st->pz.d[k];"),
         ("real pxd = (real)1 + hdt * st->px.d_h[i], pyd = (real)1 + hdt * st->py.d[j], pzd = (real)1 + hdt * st->pz.d[k];",
          "real pxd = (real)1 + hdt * st->px.d[i], pyd = (real)1 + hdt * st->py.d[j], pzd = (real)1 + hdt * st->pz.d[k];")],
        ["spml_mirror_symmetry", "spml_crank_nicolson_factors"]),
    "spml: z split dropped from the vx sum": (
        [("st->vx[p] = q[0][p] + q[1][p] + q[2][p];", "st->vx[p] = q[0][p] + q[1][p];")],
        ["spml_without_damping_is_the_plain_scheme"]),
    "spml: lambda in place of lambda + 2 mu in the sxx x split": (
        [("q[9][p] = (pxn * q[9][p] + dt * l2m * dvxdx) / pxd;", "q[9][p] = (pxn * q[9][p] + dt * lam * dvxdx) / pxd;")],
        ["spml_without_damping_is_the_plain_scheme"]),
    "spml: x and y derivative terms swapped between the vx splits": (
        [("q[0][p] = (pxn * q[0][p] + dsxxdx * dt * rho_over) / pxd;",
          "q[0][p] = (pxn * q[0][p] + dsxydy * dt * rho_over) / pxd;"),
         ("q[1][p] = (pyn * q[1][p] + dsxydy * dt * rho_over) / pyd;",
          "q[1][p] = (pyn * q[1][p] + dsxxdx * dt * rho_over) / pyd;")],
        ["spml_plane_wave_in_a_transverse_layer"]),
    "spml: wrong sign in the Crank-Nicolson denominator of the normal stresses": (
        [("real pxd = (real)1 + hdt * st->px.d[i], pyd = (real)1 + hdt * st->py.d[j], pzd = (real)1 + hdt * st->pz.d[k];\n"
          "                    q[9][p]",
          "real pxd = (real)1 - hdt * st->px.d[i], pyd = (real)1 + hdt * st->py.d[j], pzd = (real)1 + hdt * st->pz.d[k];\n"
          "                    q[9][p]")],
        ["spml_crank_nicolson_factors"]),
    "sv: force one cell off the source point": (
        [("size_t p = I2(st->isx, st->isz);\n        st->vx[p] += dt * amp / st->rho[p];",
          "size_t p = I2(st->isx + 1, st->isz);\n        st->vx[p] += dt * amp / st->rho[p];")],
        ["sv_reciprocity", "sv_mirror_symmetry"]),
    "sv: integer-point CPML coefficients at the half-point vx": (
        [("m[0][p] = st->px.b_h[i] * m[0][p] + st->px.a_h[i] * dsxxdx;",
          "m[0][p] = st->px.b[i] * m[0][p] + st->px.a[i] * dsxxdx;")],
        ["sv_mirror_symmetry"]),
    "sv: point density in place of the x average at vx": (
        [("st->vx[p] += (dsxxdx + dsxzdz) * dt / rho_half_x;", "st->vx[p] += (dsxxdx + dsxzdz) * dt / st->rho[p];")],
        ["sv_reciprocity"]),
    "rsg: sign of one rotated x derivative": (
        [("real dsxxdx = (sxxrz - sxxrx) / dx2;", "real dsxxdx = (sxxrx - sxxrz) / dx2;")],
        ["rsg_transpose_symmetry"]),
    "rsg: rotated derivatives scaled by 1/dx instead of 1/(2 dx)": (
        [("const real dx2 = (real)2 * dx, dz2 = (real)2 * dz;", "const real dx2 = dx, dz2 = dz;")],
        ["rsg_p_arrival_along_a_diagonal"]),
    "rsg: half-point CPML coefficients at the velocity nodes": (
        [("m[0][p] = st->px.b[i] * m[0][p] + st->px.a[i] * dsxxdx;\n                m[5][p]",
          "m[0][p] = st->px.b_h[i] * m[0][p] + st->px.a_h[i] * dsxxdx;\n                m[5][p]")],
        ["rsg_transpose_symmetry"]),
}


def _build_mutant(tmp_path, edits):
    src = open(os.path.join(ORACLE_DIR, "seiswave_oracle.c")).read()
    for old, new in edits:
        assert src.count(old) == 1, f"mutation site not unique or missing: {old[:60]!r}"
        src = src.replace(old, new)
    shutil.copy(os.path.join(ORACLE_DIR, "seiswave_oracle.h"), tmp_path)
    (tmp_path / "seiswave_oracle.c").write_text(src)
    lib = tmp_path / "liboracle_mutant.so"
    subprocess.run(["gcc", "-O2", "-fopenmp", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-std=c99",
                    "-DOR_REAL=double", "-shared", "-o", str(lib), str(tmp_path / "seiswave_oracle.c"), "-lm"],
                   check=True)
    return str(lib)


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc to build the mutated oracle")
@pytest.mark.parametrize("mistake", sorted(MUTANTS))
def test_pins_reject_a_mistaken_oracle(mistake, tmp_path):
    edits, pins = MUTANTS[mistake]
    lib = _build_mutant(tmp_path, edits)
    with O.library(lib, "f64"):
        for name in pins:
            with pytest.raises(AssertionError):
                getattr(P, name)()
