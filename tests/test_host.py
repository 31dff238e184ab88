"""Host-side checks that need no GPU: the C ABI library, the ctypes mirror of
the header, coefficient/PML host code and the Field wrapper."""
import os
import re

import numpy as np
import pytest
from hypothesis import given, settings
from hypothesis import strategies as st

from oracle import oracle as O
from seiswave_taichi_amd import _lib as L
from seiswave_taichi_amd.coeffs import diff_coff, diff_coff_op
from seiswave_taichi_amd.field import Field, to_device_layout
from seiswave_taichi_amd.pml import axis_pml, check_parameter, memory_slab
from tests.conftest import ROOT, gpu_available

HEADER = os.path.join(ROOT, "include", "seiswave.h")


def header_functions():
    txt = open(HEADER).read()
    return sorted(set(re.findall(r"^int(?:64_t)?\s+(sw_\w+)\s*\(", txt, flags=re.M)))


def test_library_exports_every_header_symbol():
    lib = L.load()
    names = header_functions()
    assert len(names) >= 15
    for n in names:
        assert hasattr(lib, n), n
    assert set(names) == set(L.SIGNATURES), "ctypes signature table out of sync with the header"


def test_version_and_error_channel():
    lib = L.load()
    assert lib.sw_version() == (0 << 16) | (1 << 8)
    # a null descriptor is rejected on the host, before any device work
    rc = lib.sw_ssg3d_velocity(None, None)
    assert rc == -1
    assert "null descriptor" in L.last_error()


def test_descriptor_checks_reject_bad_shapes():
    lib = L.load()
    g = L.SwSSG3D()
    g.nx = g.ny = g.nz = 5
    g.order = 2
    assert lib.sw_ssg3d_stress(g, None) == -1
    assert "too small" in L.last_error()
    g.nx = g.ny = g.nz = 32
    g.order = 9
    assert lib.sw_ssg3d_stress(g, None) == -1
    assert "order" in L.last_error()


def test_ctypes_struct_sizes_match_header_layout():
    # sw_axis: 8 float* + 2 uint8* + 2 int32 = 10*8 + 8 bytes
    assert C_sizeof(L.SwAxis) == 88
    # sw_source: 4 int32 + 6 float + pointer + int32 (+ padding)
    assert C_sizeof(L.SwSource) == 4 * 4 + 6 * 4 + 8 + 8


def C_sizeof(t):
    import ctypes
    return ctypes.sizeof(t)


@pytest.mark.parametrize("order", range(1, 9))
def test_product_coefficients_are_f32_of_oracle(order):
    np.testing.assert_array_equal(diff_coff(order), O.diff_coff(order).astype(np.float32))


def test_diff_coff_op_literals():
    # the literal constants of src/Elastic2D_CPML.py:372-389
    np.testing.assert_array_equal(diff_coff_op(2), np.float32([1.129042, -0.0430142]))
    np.testing.assert_array_equal(diff_coff_op(5)[-1], np.float32(0.0007054313))
    assert diff_coff_op(3).dtype == np.float32 and len(diff_coff_op(4)) == 4


def test_pml_parameter_validation():
    with pytest.raises(KeyError):
        check_parameter({"vp_max": 1.0}, ("pml_x_thick",))
    with pytest.raises(ValueError):
        axis_pml(40, 2, 10.0, 1e-3, 0.0, 3000.0, 1e-6, 30.0, 2.0, 1.0, True, True)


def test_memory_slab_whole_axis_when_pml_covers_it():
    n = 12
    ax = axis_pml(n, 2, 10.0, 1e-3, 60.0, 3000.0, 1e-6, 30.0, 2.0, 1.0, True, True)
    assert ax.lo + (n - ax.hi) <= n
    nontriv = (ax.a != 0) | (ax.k != 1) | (ax.a_h != 0) | (ax.k_h != 1)
    assert not nontriv[ax.lo:ax.hi].any()


def test_memory_slab_refuses_coefficients_outside_the_pml_test():
    a = np.zeros(10, np.float32)
    k = np.ones(10, np.float32)
    a[5] = -0.1
    inside = np.zeros(10, np.uint8)
    with pytest.raises(ValueError):
        memory_slab(a, k, a.copy(), k.copy(), inside, inside)


def test_field_presents_reference_index_order():
    ref = np.arange(2 * 3 * 4, dtype=np.float32).reshape(2, 3, 4)
    t = to_device_layout(ref, "cpu")
    assert t.shape == (4, 3, 2) and t.is_contiguous()
    f = Field(t)
    assert f.shape == (2, 3, 4)
    np.testing.assert_array_equal(f.to_numpy(), ref)
    assert f[1, 2, 3] == ref[1, 2, 3]
    f[0, 1, 2] = -5.0
    assert t[2, 1, 0].item() == -5.0
    f.from_numpy(ref * 2)
    np.testing.assert_array_equal(np.asarray(f), ref * 2)


@pytest.mark.skipif(gpu_available(), reason="checks the no-GPU error path")
def test_solver_refuses_to_run_without_gpu():
    from seiswave_taichi_amd.Elastic3D_SSG_CPML import ElasticWAVE
    v = np.ones((16, 16, 16), np.float32)
    with pytest.raises(RuntimeError, match="GPU"):
        ElasticWAVE(v, v, v, 1.0, 1.0, 1.0, 1e-4, 5, 5, 5, 6, 6, 6, 10, 2)


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "seiswave_taichi_amd")
    for dirpath, _, files in os.walk(pkg):
        for fn in files:
            if fn.endswith((".py", ".hip", ".h", ".cpp")):
                txt = open(os.path.join(dirpath, fn)).read()
                assert "import oracle" not in txt and "from oracle" not in txt, fn
                assert "seiswave_oracle" not in txt, fn


def test_install_src_alias():
    import sys
    import seiswave_taichi_amd as pkg
    def ours(k):
        return k in ("src", "taichi") or k.startswith("src.")
    had_taichi = "taichi" in sys.modules
    saved = {k: sys.modules.pop(k) for k in list(sys.modules) if ours(k)}
    try:
        pkg.install_src_alias()
        from src.Elastic3D_SSG_Tensor import ElasticWAVE  # noqa: F401
        from src.BaseKernel import Ricker  # noqa: F401  (Relay_wave.py:5)
        from src.wigb import wigb  # noqa: F401  (test3d_homegenous_tensor.py:5)
        import src.Elastic2D_CPML as m
        assert m.__name__ == "seiswave_taichi_amd.Elastic2D_CPML"
        import taichi as ti
        if not had_taichi:
            assert ti.__name__ == "seiswave_taichi_amd.taichi_shim"
    finally:
        for k in [k for k in sys.modules if ours(k)]:
            sys.modules.pop(k)
        sys.modules.update(saved)


def test_base_kernel_ricker_closed_form():
    """src/BaseKernel.py:4-12: -s*2*a*tau*exp(-a*tau^2), tau = n*dt - 1/f; zero at
    tau = 0, odd about it, extrema at tau = -+1/sqrt(2a) with value
    +-s*sqrt(2a)*exp(-1/2)."""
    from seiswave_taichi_amd.BaseKernel import Ricker, Ricker2
    f, dt, s = 25.0, 1e-4, 1000.0
    a = np.pi ** 2 * f ** 2
    for n in (0, 100, 350, 400, 450, 1999):
        tau = n * dt - 1 / f
        assert Ricker(n, dt, f, s) == pytest.approx(-s * 2 * a * tau * np.exp(-a * tau ** 2), rel=1e-5, abs=1e-3)
    assert abs(Ricker(400, dt, f, s)) < 1e-2 * s
    assert Ricker(300, dt, f, s) == pytest.approx(-Ricker(500, dt, f, s), rel=1e-4)
    n_peak = (1 / f - 1 / np.sqrt(2 * a)) / dt
    peak = max(Ricker(n, dt, f, s) for n in range(int(n_peak) - 2, int(n_peak) + 3))
    assert peak == pytest.approx(s * np.sqrt(2 * a) * np.exp(-0.5), rel=1e-3)
    assert Ricker2(400, dt, 0.04, f, s) == pytest.approx(s, rel=1e-6)  # tau = 0: the peak


def test_wigb_draws_one_wiggle_per_trace():
    import matplotlib
    matplotlib.use("Agg")
    import matplotlib.pyplot as plt
    from seiswave_taichi_amd.wigb import wigb
    t = np.linspace(0, 1, 200)[:, None]
    a = np.sin(40 * t) * np.exp(-((t - 0.5) ** 2) / 0.01) * np.arange(1, 6)[None]
    a[:, 2] = 0.0
    fig = wigb(a)
    ax = fig.axes[0]
    assert len(ax.lines) == 5 and len(ax.collections) == 4  # the zero trace is a bare line
    assert ax.yaxis_inverted()
    plt.close(fig)


def _fake_3d(n=24, order=2):
    """A descriptor that passes the host-side checks without any device memory
    (the pointers are never dereferenced before the rejected argument)."""
    import ctypes as C
    g = L.SwSSG3D()
    g.nx = g.ny = g.nz = n
    g.order, g.cls = order, L.SW_CLASS_TENSOR
    fake = C.cast(C.c_void_p(0x1000), L._fp)
    for name in ("rho", "lam", "mu", "vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz",
                 "mem_x", "mem_y", "mem_z"):
        setattr(g, name, fake)
    for ax in (g.px, g.py, g.pz):
        for f in ("a", "b", "k", "a_h", "b_h", "k_h", "d", "d_h"):
            setattr(ax, f, fake)
        ax.lo, ax.hi = 6, n - 6
    return g


def test_run_flags_and_free_surface_arguments_are_checked():
    import ctypes as C
    lib = L.load()
    g = _fake_3d()
    src = L.SwSource()
    src.kind = L.SW_SRC_NONE
    assert lib.sw_ssg3d_run_ex(C.byref(g), C.byref(src), None, 0, 4, 0x10, None) == -1
    assert "unknown run flags" in L.last_error()
    g.order = g.nz  # the free-surface plane must lie inside the grid
    assert lib.sw_ssg3d_free_surface(C.byref(g), None) == -1


def test_force_adjoint_needs_its_accumulator():
    import ctypes as C
    lib = L.load()
    nx, nz = 30, 26
    P = nx * nz
    base = 0x100000
    g = L.SwSSG2D()
    g.nx, g.nz, g.order, g.scheme = nx, nz, 2, L.SW_2D_VS
    fp = lambda a: C.cast(C.c_void_p(a), L._fp)  # noqa: E731
    g.rho = g.lam = g.mu = fp(base)
    for r, name in enumerate(("vx", "vz", "sxx", "szz", "sxz")):
        setattr(g, name, fp(base + 4 * r * P))
    g.mem = fp(base + 4 * 5 * P)
    for ax in (g.px, g.pz):
        for f in ("a", "b", "k", "a_h", "b_h", "k_h", "d", "d_h"):
            setattr(ax, f, fp(base))
        ax.in_ = ax.in_h = C.cast(C.c_void_p(base), L._u8p)
        ax.lo, ax.hi = 0, nx
    g.pz.hi = nz
    q = L.SwAdj2D()
    for name in ("vx", "vz", "sxx", "szz", "sxz", "mem", "tmp", "g_b1", "g_b4", "g_lamh", "g_muh", "g_muz"):
        setattr(q, name, fp(base))
    src = L.SwSource()
    src.kind, src.ix, src.iz = L.SW_SRC_FORCE_VZ, 10, 10
    src.amp, src.namp = fp(base), 100
    rec = L.SwReceivers()
    rec.nrec, rec.ncomp, rec.nt = 0, 1, 100
    rc = lib.sw_ssg2d_adjoint_segment(C.byref(g), C.byref(src), C.byref(q), C.byref(rec), fp(base), 0, 5, None)
    assert rc == -1 and "g_src" in L.last_error()


def test_header_struct_layouts_match_ctypes(tmp_path):
    """sizeof and field offsets of every public struct, from the C compiler on
    include/seiswave.h, against the ctypes mirror."""
    import ctypes as C
    import shutil
    import subprocess
    if shutil.which("gcc") is None:
        pytest.skip("no C compiler")
    structs = {"sw_axis": L.SwAxis, "sw_ssg3d": L.SwSSG3D, "sw_ssg3d_shadow": None, "sw_ssg2d": L.SwSSG2D,
               "sw_adj2d": L.SwAdj2D, "sw_source": L.SwSource, "sw_receivers": L.SwReceivers, "sw_band": L.SwBand}
    structs = {k: v for k, v in structs.items() if v is not None}
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "seiswave.h"', "int main(void) {"]
    for cname, ct in structs.items():
        lines.append(f'printf("{cname} %zu\\n", sizeof({cname}));')
        for fname, _ in ct._fields_:
            cf = "in" if fname == "in_" else fname
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {cf}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = dict(line.rsplit(" ", 1) for line in subprocess.run([str(exe)], capture_output=True, text=True,
                                                               check=True).stdout.splitlines())
    for cname, ct in structs.items():
        assert int(out[cname]) == C.sizeof(ct), cname
        for fname, _ in ct._fields_:
            assert int(out[f"{cname}.{fname}"]) == getattr(ct, fname).offset, f"{cname}.{fname}"


def test_taichi_shim_covers_the_driver_calls():
    """The Taichi calls the reference drivers make outside the solver classes
    (test3d_homegenous_tensor.py:8-58): init, fields with from_numpy and item
    assignment, int-typed receiver fields, sqrt, sync."""
    from seiswave_taichi_amd import taichi_shim as ti
    ti.init(arch=ti.vulkan)
    vp = ti.field(dtype=ti.f32, shape=(4, 5, 6))
    vp.from_numpy(np.full((4, 5, 6), 3000.0))
    assert vp.shape == (4, 5, 6) and vp.to_numpy().dtype == np.float32 and vp[1, 2, 3] == 3000.0
    rsx = ti.field(dtype=int, shape=(1, 81))
    rsx.from_numpy(np.linspace(21, 101, 81, dtype=int)[None])
    assert rsx.to_numpy().dtype == np.int32 and rsx[0, 80] == 101
    mt = ti.field(dtype=ti.f32, shape=(3, 3))
    mt[1, 2] = -1 / ti.sqrt(2)
    assert mt[1, 2] == pytest.approx(-0.70710677)
    with pytest.raises(ValueError):
        vp.from_numpy(np.zeros((4, 5)))
    ti.sync()


def test_grid_size_limits_are_checked():
    """The 3-D sweeps address each z chunk with 32-bit offsets and receivers
    with int32 cell indices: grids beyond those ranges are refused with an
    error before anything is launched (no silent wrap-around)."""
    import ctypes as C
    lib = L.load()
    src = L.SwSource()
    src.kind = L.SW_SRC_NONE

    def run(nx, ny, nz, lo=6):
        g = _fake_3d()
        g.nx, g.ny, g.nz = nx, ny, nz
        for ax, n in ((g.px, nx), (g.py, ny), (g.pz, nz)):
            ax.lo, ax.hi = lo, n - lo
        rc = lib.sw_ssg3d_run_ex(C.byref(g), C.byref(src), None, 0, 1, 0x10, None)
        return rc, L.last_error()

    rc, msg = run(1300, 1300, 1300)  # 2.2e9 cells
    assert rc == -1 and "int32 cell index" in msg
    rc, msg = run(12000, 12000, 8)  # 144e6-cell planes: a 3-plane chunk + halo exceeds 4 GiB
    assert rc == -1 and "z chunk" in msg
    rc, msg = run(200, 3000, 3000, lo=400)  # 800-cell x slabs: 24*800*3000*3000 B
    assert rc == -1 and "memory slabs" in msg
    # 1100 x 1000 x 1000 (4.4 GB arrays) passes the size checks and fails only
    # on the next one (the unknown run flag)
    rc, msg = run(1100, 1000, 1000)
    assert rc == -1 and "unknown run flags" in msg


def test_sweep_entry_points_refuse_the_other_class():
    """Each 3-D sweep entry point validates the arrays of the class it runs:
    the split-field sweeps refuse a CPML/Tensor descriptor (whose split arrays
    were never checked) and the CPML sweeps a split-field one; an unknown
    class is refused everywhere.  The checks return before any launch."""
    import ctypes as C
    lib = L.load()
    g = _fake_3d()
    for cls in (L.SW_CLASS_CPML, L.SW_CLASS_TENSOR):
        g.cls = cls
        for fn in (lib.sw_ssg3d_velocity_spml, lib.sw_ssg3d_stress_spml):
            assert fn(C.byref(g), None) == -1
            assert "split-field descriptor" in L.last_error()
    g.cls = L.SW_CLASS_SPML
    g.split = C.cast(C.c_void_p(0x1000), L._fp)
    for ax in (g.px, g.py, g.pz):
        ax.in_ = ax.in_h = C.cast(C.c_void_p(0x1000), L._u8p)
    for fn in (lib.sw_ssg3d_velocity, lib.sw_ssg3d_stress):
        assert fn(C.byref(g), None) == -1
        assert "CPML/Tensor descriptor" in L.last_error()
    # a split-field descriptor without split arrays fails the split-field check
    g.split = None
    assert lib.sw_ssg3d_velocity_spml(C.byref(g), None) == -1
    assert "split-field arrays missing" in L.last_error()
    g.cls = 7
    for fn in (lib.sw_ssg3d_velocity, lib.sw_ssg3d_stress, lib.sw_ssg3d_velocity_spml, lib.sw_ssg3d_stress_spml):
        assert fn(C.byref(g), None) == -1
        assert "unknown 3-D class" in L.last_error()


def test_launch_plan_rule():
    """Automatic plan: five launches (x; x, y twice; x, y, z twice) when the
    CPML-free bulk is >= 75 % of the updated cells and every slab launch has
    at least as many blocks as the device has compute units (256 where no
    device answers), else one launch with CPML on every axis; the two forced
    modes; an unknown mode is refused and the mode is process-wide."""
    lib = L.load()
    g = _fake_3d(n=60)
    for ax in (g.px, g.py, g.pz):
        ax.lo, ax.hi = 3, 57      # thin slabs: bulk (53/56)^2 = 0.90, slab launches of 13-16 blocks
    assert L.launch_plan(g) == [7]
    big = _fake_3d(n=400)
    for ax in (big.px, big.py, big.pz):
        ax.lo, ax.hi = 10, 390    # bulk (380/396)^2 = 0.92, slab launches of 1,200+ blocks
    assert L.launch_plan(big) == [1, 3, 3, 7, 7]
    mid = _fake_3d(n=360)
    for ax in (mid.px, mid.py, mid.pz):
        ax.lo, ax.hi = 3, 357     # 1-plane z slabs (about 500 blocks), 1-row y slabs (2 column blocks x 118 chunks)
    assert L.launch_plan(mid) == [7]
    thin = _fake_3d(n=168)        # the learned boundary's 4-cell CPML grid of the 200^3 case
    for ax in (thin.px, thin.py, thin.pz):
        ax.lo, ax.hi = 6, 161     # bulk (155/164)^2 = 0.89, z slabs of about 110 blocks
    assert L.launch_plan(thin) == [7]
    for ax in (g.px, g.py, g.pz):
        ax.lo, ax.hi = 10, 50     # bulk (40/56)^2 = 0.51
    assert L.launch_plan(g) == [7]
    prev = lib.sw_set_launch_plan(2)  # SW_PLAN_BOXES
    try:
        assert prev == 0
        assert L.launch_plan(g) == [1, 3, 3, 7, 7]
        assert lib.sw_set_launch_plan(1) == 2  # SW_PLAN_ONE
        assert L.launch_plan(big) == [7]
        assert lib.sw_set_launch_plan(3) == -1
        assert "unknown launch plan mode" in L.last_error()
    finally:
        lib.sw_set_launch_plan(0)
    assert L.launch_plan(g) == [7]


def test_src_package_at_repository_root():
    """``from src.<Module> import ElasticWAVE`` (test3d_homegenous_tensor.py:4,
    test2d_homegenous.py:4, ...) resolves, in a fresh interpreter with only the
    repository root on sys.path, to this package's classes."""
    import subprocess
    import sys
    code = ("import src.Elastic2D_CPML as a, src.Elastic2D_sparse as b, src.Elastic3D_SSG_CPML as c, "
            "src.Elastic3D_SSG_SPML as d, src.Elastic3D_SSG_Tensor as e\n"
            "from src.wigb import *\nfrom src.BaseKernel import Ricker, Ricker2\n"
            "print(' '.join(m.ElasticWAVE.__module__ for m in (a, b, c, d, e)), wigb.__module__)")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True, timeout=300,
                       env=dict(os.environ, PYTHONPATH=ROOT))
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.split() == [f"seiswave_taichi_amd.{m}" for m in (
        "Elastic2D_CPML", "Elastic2D_sparse", "Elastic3D_SSG_CPML", "Elastic3D_SSG_SPML",
        "Elastic3D_SSG_Tensor", "wigb")]



@pytest.mark.parametrize("case", [
    # (shape, order, (xlo, xhi), (ylo, yhi), (zlo, zhi))
    ((40, 36, 44), 2, (9, 31), (8, 28), (10, 34)),
    ((30, 30, 30), 1, (0, 24), (5, 30), (6, 23)),      # x-lo and y-hi faces without PML
    ((33, 28, 31), 4, (0, 0), (9, 19), (10, 21)),      # no split-free x range: every row is shell
    ((26, 29, 24), 3, (8, 18), (9, 20), (0, 0)),       # no split-free z range: every plane is shell
    ((24, 24, 24), 2, (0, 0), (0, 0), (0, 0)),         # all shell
    ((25, 27, 23), 2, (0, 25), (0, 27), (0, 23)),      # no PML at all: empty shell
])
def test_split_shell_layout_matches_the_sweeps(case):
    """The split-field class keeps its 24 split components only on the shell
    around the split-free box.  The (cell, slot) pairs the shell launch
    computes (the kernels' own index function, run on the host) cover every
    updated cell outside the box exactly once, with distinct slots inside
    each component's storage, at the slots the host-side unpacking reads."""
    _check_split_shell_layout(case)


@settings(max_examples=25, deadline=None)
@given(st.data())
def test_split_shell_layout_random_boxes(data):
    """The same checks on random grid shapes, orders and split-free boxes."""
    order = data.draw(st.integers(1, 4))
    shape, ranges = [], []
    for _ in range(3):
        n = data.draw(st.integers(2 * order + 2, 24))
        shape.append(n)
        if data.draw(st.booleans()):
            ranges.append((0, 0))  # no split-free range on this axis: all shell
        else:
            lo = data.draw(st.integers(0, n - 1))
            ranges.append((lo, data.draw(st.integers(lo + 1, n))))
    _check_split_shell_layout((tuple(shape), order, *ranges))


def _check_split_shell_layout(case):
    from seiswave_taichi_amd._engine3d import SplitShell
    (nx, ny, nz), order, rx, ry, rz = case
    g = _fake_3d(order=order)
    g.cls = L.SW_CLASS_SPML
    g.nx, g.ny, g.nz = nx, ny, nz
    for ax, (lo, hi) in zip((g.px, g.py, g.pz), (rx, ry, rz)):
        ax.lo, ax.hi = lo, hi
    sh = SplitShell((nx, ny, nz), (rx, ry, rz))
    assert L.load().sw_ssg3d_split_cells(g) == sh.cells
    cell, slot = L.split_map(g, sweep=1)
    # the velocity sweep's launches update the same cells at the same slots
    cv, sv = L.split_map(g, sweep=0)
    by_cell = np.argsort(cv)
    np.testing.assert_array_equal(cv[by_cell], np.sort(cell))
    np.testing.assert_array_equal(sv[by_cell], slot[np.argsort(cell)])
    N = order
    k, r = np.divmod(cell, nx * ny)
    j, i = np.divmod(r, nx)
    upd = (slice(N, nz - N), slice(N, ny - N), slice(N, nx - N))
    inside = np.zeros((nz, ny, nx), bool)
    inside[rz[0]:rz[1], ry[0]:ry[1], rx[0]:rx[1]] = True
    want = np.zeros((nz, ny, nx), bool)
    want[upd] = ~inside[upd]
    got = np.zeros((nz, ny, nx), np.int64)
    np.add.at(got, (k, j, i), 1)
    np.testing.assert_array_equal(got, want.astype(np.int64))
    assert np.unique(slot).size == slot.size
    assert slot.size == 0 or (slot.min() >= 0 and slot.max() < sh.cells)
    expect = np.array([sh.slot(a, b, c) for a, b, c in zip(i, j, k)], np.int64)
    np.testing.assert_array_equal(slot, expect)
    # unpacking puts slot s of a component at its cell
    comp = np.zeros(sh.cells, np.float64)
    comp[slot] = cell + 1.0
    full = sh.unpack(comp).reshape(-1)
    np.testing.assert_array_equal(full[cell], cell + 1.0)
    assert np.count_nonzero(full) == cell.size
    # the 24 components of the blocked storage occupy distinct floats, all of them
    idx = np.concatenate([sh.index(c, np.arange(sh.cells)) for c in range(24)])
    assert np.unique(idx).size == idx.size == 24 * sh.cells
    assert idx.size == 0 or (idx.min() >= 0 and idx.max() < sh.floats)
    if sh.cells % 64 == 0:
        assert idx.size == sh.floats
    storage = np.zeros(sh.floats)
    storage[sh.index(13, slot)] = cell + 2.0
    np.testing.assert_array_equal(sh.component(storage, 13)[slot], cell + 2.0)


def test_split_shell_refuses_bad_ranges():
    g = _fake_3d()
    g.cls = L.SW_CLASS_SPML
    g.px.lo, g.px.hi = 10, 5
    assert L.load().sw_ssg3d_split_cells(g) == -1
    assert "split-free range" in L.last_error()
