"""Model and SEG-Y I/O (CPU): formulas of the overthrust driver, IBM floats
against published bit patterns, and SEG-Y round trips with the shot grouping
of ReadSegy.read_segy."""
import numpy as np
import pytest

from seiswave_taichi_amd import io


def test_mat_model_follows_the_overthrust_driver(tmp_path):
    """vs = vp/1.7 and Gardner rho = 1.741*(vp/1000)**0.25*1000
    (test3d_homegenous_overhust.py:9-15)."""
    vp = np.random.default_rng(0).uniform(1500, 5500, (6, 5, 4))
    p = tmp_path / "model.mat"
    io.save_mat_model(p, vp)
    vs, vp2, rho = io.load_mat_model(p)
    assert vp2.dtype == np.float32 and vp2.shape == (6, 5, 4)
    np.testing.assert_array_equal(vp2, vp.astype(np.float32))
    np.testing.assert_allclose(vs, vp / 1.7, rtol=1e-6)
    np.testing.assert_allclose(rho, 1741.0 * (vp / 1000.0) ** 0.25, rtol=1e-6)
    # 1000 m/s maps to 1741 kg/m^3 exactly
    io.save_mat_model(p, np.full((2, 2, 2), 1000.0))
    assert float(io.load_mat_model(p)[2][0, 0, 0]) == pytest.approx(1741.0)
    with pytest.raises(KeyError):
        io.load_mat_model(p, key="Vs")


def test_ibm_floats_match_published_patterns():
    # -118.625 = 0xC276A000 (the textbook example); 100 = 0x42640000; 1 = 0x41100000;
    # 0.1 = 0x4019999A rounded (0x40199999 truncated)
    words = [0xC276A000, 0x42640000, 0x41100000, 0x4019999A, 0x00000000]
    # 0x4019999A holds 0x19999A / 2**24 exactly (0.1 to 24 fraction bits)
    want = np.float32([-118.625, 100.0, 1.0, 0x19999A / 2.0 ** 24, 0.0])
    np.testing.assert_array_equal(io.ibm_to_ieee(words), want)
    np.testing.assert_array_equal(io.ieee_to_ibm([-118.625, 100.0, 1.0, 0.1, 0.0]), np.uint32(words))
    # normalisation: the leading hex digit of the fraction is never 0
    w = io.ieee_to_ibm(np.float32([3.0e-5, 7.7e6, 1.0 / 16, 15.999]))
    assert np.all((w & 0x00F00000) != 0)


def test_ibm_round_trip_precision():
    rng = np.random.default_rng(1)
    x = (rng.standard_normal(20000) * 10.0 ** rng.integers(-30, 30, 20000)).astype(np.float32)
    y = io.ibm_to_ieee(io.ieee_to_ibm(x))
    # 24-bit fraction with a base-16 exponent: at least 21 significant bits
    assert np.max(np.abs(y - x) / np.abs(x)) <= 2.0 ** -21


@pytest.mark.parametrize("fmt", [1, 5])
def test_segy_round_trip_and_shot_grouping(tmp_path, fmt):
    rng = np.random.default_rng(2)
    nshot, ntr, ns = 3, 7, 50
    gathers = rng.standard_normal((nshot, ns, ntr)).astype(np.float32)
    traces = np.concatenate([g.T for g in gathers])           # shot-major trace order
    sx = np.repeat([100, 250, 400], ntr)
    p = tmp_path / "shots.sgy"
    io.write_segy(p, traces, sx, dt=0.002, fmt=fmt)
    samples, source_x, dt = io.read_segy_traces(p)
    np.testing.assert_array_equal(source_x, sx)
    assert dt == pytest.approx(0.002)
    data = io.read_segy(p)
    assert data.shape == (nshot, ns, ntr)
    if fmt == 5:
        np.testing.assert_array_equal(data, gathers)
    else:
        np.testing.assert_allclose(data, gathers, rtol=2.0 ** -21, atol=0)
    # an explicit shot count splits the traces evenly, as the reference does
    assert io.read_segy(p, shotnum=1).shape == (ns, nshot * ntr)


def test_segy_rejects_truncated_files(tmp_path):
    p = tmp_path / "bad.sgy"
    io.write_segy(p, np.zeros((2, 10)), [1, 1], 0.001)
    raw = p.read_bytes()
    p.write_bytes(raw[:-3])
    with pytest.raises(ValueError):
        io.read_segy_traces(p)
