"""Seeded heterogeneous model generator (the full-scale 3-D workload)."""
import numpy as np

from seiswave_taichi_amd.models import cfl_dt, gaussian_noise_field, layered_folded_model


def test_model_is_deterministic_and_physical():
    shape = (64, 48, 40)
    a = layered_folded_model(shape, 12.5, seed=3)
    b = layered_folded_model(shape, 12.5, seed=3)
    c = layered_folded_model(shape, 12.5, seed=4)
    for x, y in zip(a, b):
        np.testing.assert_array_equal(x, y)
    assert not np.array_equal(a[1], c[1])
    vs, vp, rho = a
    water = vs == 0
    assert water[:, :, 0].all() and not water[:, :, -1].any()
    np.testing.assert_array_equal(vp[water], np.float32(1500.0))
    np.testing.assert_allclose(vs[~water], vp[~water] / 1.73, rtol=1e-6)
    np.testing.assert_allclose(rho, 1741.0 * (vp / 1000.0) ** 0.25, rtol=1e-6)
    # velocity rises with depth on average
    assert vp[:, :, -1].mean() > vp[:, :, vp.shape[2] // 2].mean() > vp[:, :, 1].mean()


def test_perturbation_statistics():
    rng = np.random.default_rng(0)
    corr = 4.8
    f = gaussian_noise_field((96, 96, 96), corr, rng)
    assert abs(f.mean()) < 1e-6 and abs(f.std() - 1.0) < 1e-5
    # Gaussian kernel of std s gives a field autocorrelation exp(-r^2/(4 s^2))
    r = 5
    ac = np.mean(f[:-r] * f[r:])
    expect = np.exp(-r ** 2 / (4 * corr ** 2))
    assert abs(ac - expect) < 0.12, (ac, expect)


def test_cfl_dt():
    dt = cfl_dt(5500.0, (12.5, 12.5, 12.5), 0.45)
    assert abs(5500.0 * dt * np.sqrt(3 / 12.5 ** 2) - 0.45) < 1e-12


def test_2d_anomaly_model_is_seeded_and_bounded():
    from seiswave_taichi_amd.models import layered_2d_with_anomalies
    (vp0, vs0, rho), (vp1, vs1, rho1) = layered_2d_with_anomalies((200, 90), 10.0, seed=3)
    (wp0, _, _), (wp1, _, _) = layered_2d_with_anomalies((200, 90), 10.0, seed=3)
    np.testing.assert_array_equal(vp1, wp1)
    np.testing.assert_array_equal(vp0, wp0)
    r = vp1 / vp0
    assert 1.095 < r.max() <= 1.1 + 1e-6 and 0.9 - 1e-6 <= r.min() < 0.905
    np.testing.assert_allclose(vs1 / vs0, r, rtol=1e-6)
    np.testing.assert_allclose(vp0 / vs0, 1.73, rtol=1e-6)
    np.testing.assert_array_equal(rho, rho1)
    # background: vp non-decreasing with depth in every column
    assert np.all(np.diff(vp0, axis=1) >= 0)
