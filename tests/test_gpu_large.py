"""Arrays above 4 GiB (1100 x 1000 x 1000 cells, 4.4 GB per array, about 57 GB
of device memory): the sweeps rebase their pointers per block there.  A size-
independent check: in a homogeneous medium, a moment source far from every
boundary leaves every cell beyond its light cone (2N cells per step) exactly
zero, and the cone itself evolves bit for bit as on a small grid with the same
local setup.  The source sits at plane 975, so the cone straddles the 4 GiB
byte offset (plane 976.2) of every full-grid array."""
import numpy as np
import pytest
import torch

from tests.cases import pml_dict, unit_mt

pytestmark = pytest.mark.gpu

STEPS, ORDER, NPML, D, DT, F = 3, 2, 5, 10.0, 1e-3, 10.0
R = 2 * ORDER * STEPS + 1  # light-cone radius in cells


def _run(shape, src, rec):
    from seiswave_taichi_amd.Elastic3D_SSG_Tensor import ElasticWAVE
    vs, vp, rho = (np.full(shape, v, np.float32) for v in (2000.0, 3500.0, 2200.0))
    m = unit_mt(4)
    mt3 = np.zeros((3, 3), np.float32)
    mt3[0, 0], mt3[1, 1], mt3[2, 2], mt3[0, 1], mt3[0, 2], mt3[1, 2] = m
    w = ElasticWAVE(vs, vp, rho, D, D, D, DT, *src, rec[:, 0][None], rec[:, 1][None], rec[:, 2][None], mt3, 1.0,
                    STEPS, ORDER, F, components=("vx", "vz"))
    del vs, vp, rho
    w.SetADEPML3D([True] * 6, pml_dict(3500.0, NPML, (D, D, D), F, 2000.0))
    w.run(STEPS)
    torch.cuda.synchronize()
    return w


def test_arrays_above_4gib_match_small_grid():
    props = torch.cuda.get_device_properties(0)
    if props.total_memory < 100 * 2 ** 30:
        pytest.skip("needs a GPU with more than 100 GiB")
    big, small = (1100, 1000, 1000), (48, 48, 48)
    sb, ss = (550, 500, 975), (24, 24, 24)
    off = np.array([[0, 0, 0], [3, -2, 4], [-5, 6, -7], [8, 1, 9]])
    wb = _run(big, sb, np.array(sb) + off)
    assert 4 * np.prod(big, dtype=np.int64) > 2 ** 32
    fields = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    got, outside = {}, {}
    for name in fields:
        t = wb._e.f[name]  # device layout [z][y][x]
        box = t[sb[2] - R:sb[2] + R + 1, sb[1] - R:sb[1] + R + 1, sb[0] - R:sb[0] + R + 1]
        got[name] = box.cpu().numpy()
        outside[name] = int(torch.count_nonzero(t)) - int(torch.count_nonzero(box))
    data_b = wb.data.to_numpy()
    del wb
    torch.cuda.empty_cache()
    ws = _run(small, ss, np.array(ss) + off)
    for name in fields:
        t = ws._e.f[name]
        ref = t[ss[2] - R:ss[2] + R + 1, ss[1] - R:ss[1] + R + 1, ss[0] - R:ss[0] + R + 1].cpu().numpy()
        assert np.abs(ref).max() > 0, name
        np.testing.assert_array_equal(got[name], ref, err_msg=name)
        assert outside[name] == 0, (name, outside[name])
    np.testing.assert_array_equal(data_b, ws.data.to_numpy())


def _run_spml(shape, src, rec):
    from seiswave_taichi_amd.Elastic3D_SSG_SPML import ElasticWAVE
    vs, vp, rho = (np.full(shape, v, np.float32) for v in (2000.0, 3500.0, 2200.0))
    w = ElasticWAVE(vs, vp, rho, D, D, D, DT, *src, *rec, STEPS, ORDER, F, 3.0)
    del vs, vp, rho
    w.SetADEPML3D([True] * 6, pml_dict(3500.0, NPML, (D, D, D), F, 2000.0))
    w.run(STEPS)
    torch.cuda.synchronize()
    return w


def _cone_splits(w, c):
    """Split values of the shell cells in the light cone around cell c, per
    component, read from the shell storage on the device."""
    sh = w._e.shell
    cells = [(i, j, k) for k in range(c[2] - R, c[2] + R + 1) for j in range(c[1] - R, c[1] + R + 1)
             for i in range(c[0] - R, c[0] + R + 1)
             if not (sh.xlo <= i < sh.xhi and sh.ylo <= j < sh.yhi and sh.zlo <= k < sh.zhi)]
    slots = np.array([sh.slot(*p) for p in cells], np.int64)
    idx = np.stack([sh.index(comp, slots) for comp in range(24)])
    return w._e.split[torch.from_numpy(idx).to(w._e.split.device)].cpu().numpy(), len(cells)


def test_split_field_arrays_above_4gib_match_small_grid():
    """The split-field class on the 1100 x 1000 x 1000 grid (4.4 GB per field
    array): a force source 15 planes below the z-hi face, so that its light
    cone straddles the 4 GiB byte offset (plane 976.2) and reaches into the
    z-hi PML slab, evolves bit for bit as on a 48^3 grid with the source at
    the same distance from its z-hi face, split values included; nothing
    outside the cone moves."""
    props = torch.cuda.get_device_properties(0)
    if props.total_memory < 100 * 2 ** 30:
        pytest.skip("needs a GPU with more than 100 GiB")
    big, small = (1100, 1000, 1000), (48, 48, 48)
    sb, ss = (550, 500, 985), (24, 24, 33)
    wb = _run_spml(big, sb, np.array(sb) + [2, -3, 4])
    assert 4 * np.prod(big, dtype=np.int64) > 2 ** 32
    assert wb._e.shell.zhi - sb[2] == 7 and sb[2] + R > wb._e.shell.zhi  # the cone reaches the slab
    fields = ("vx", "vy", "vz", "sxx", "syy", "szz", "sxy", "sxz", "syz")
    got, outside = {}, {}
    for name in fields:
        t = wb._e.f[name]
        box = t[sb[2] - R:sb[2] + R + 1, sb[1] - R:sb[1] + R + 1, sb[0] - R:sb[0] + R + 1]
        got[name] = box.cpu().numpy()
        outside[name] = int(torch.count_nonzero(t)) - int(torch.count_nonzero(box))
    split_b, n_shell = _cone_splits(wb, sb)
    split_nonzero = int(torch.count_nonzero(wb._e.split))
    data_b = wb.data.to_numpy()
    del wb
    torch.cuda.empty_cache()
    ws = _run_spml(small, ss, np.array(ss) + [2, -3, 4])
    assert ws._e.shell.zhi - ss[2] == 7
    for name in fields:
        t = ws._e.f[name]
        ref = t[ss[2] - R:ss[2] + R + 1, ss[1] - R:ss[1] + R + 1, ss[0] - R:ss[0] + R + 1].cpu().numpy()
        assert np.abs(ref).max() > 0, name
        np.testing.assert_array_equal(got[name], ref, err_msg=name)
        assert outside[name] == 0, (name, outside[name])
    split_s, n_small = _cone_splits(ws, ss)
    assert n_shell == n_small > 0
    np.testing.assert_array_equal(split_b, split_s)
    assert np.count_nonzero(split_s) > 0 and split_nonzero == np.count_nonzero(split_b)
    np.testing.assert_array_equal(data_b, ws.data.to_numpy())
