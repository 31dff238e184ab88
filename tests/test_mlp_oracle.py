"""Pins of the MLP oracle itself (no GPU): finite differences of its loss."""
import numpy as np

from oracle import mlp_ref as M


def test_mlp_reference_gradient_matches_finite_differences():
    rng = np.random.default_rng(0)
    n = M.NFEAT * M.NH + M.NH + M.NH * M.NH + M.NH + M.NH * M.NOUT + M.NOUT
    p = np.concatenate([0.2 * rng.standard_normal(n), rng.uniform(0.5, 2.0, M.NFEAT),
                        rng.uniform(0.5, 2.0, M.NOUT)])
    X = rng.standard_normal((16, M.NFEAT))
    T = rng.standard_normal((16, M.NOUT))
    L, g = M.loss_and_grad(p, X, T)
    for idx in rng.choice(n, 40, replace=False):
        h = 1e-6
        pp, pm = p.copy(), p.copy()
        pp[idx] += h
        pm[idx] -= h
        fd = (M.loss_and_grad(pp, X, T)[0] - M.loss_and_grad(pm, X, T)[0]) / (2 * h)
        assert abs(fd - g[idx]) <= 1e-6 * max(1.0, abs(fd)) + 1e-9, idx


def test_adam_first_step_is_lr_times_sign():
    p = np.zeros(3)
    g = np.array([2.0, -0.5, 1e-3])
    p1, _, _ = M.adam(p, g, np.zeros(3), np.zeros(3), 0.01, 0.9, 0.999, 1e-12, 1)
    np.testing.assert_allclose(p1, -0.01 * np.sign(g), rtol=1e-6)


def test_correction_is_positively_homogeneous_and_zero_on_silence():
    """The per-column scale makes the correction proportional to the band:
    c(a X) = a c(X) for a > 0, and a silent column gets none."""
    rng = np.random.default_rng(5)
    p = rng.standard_normal(11280) * 0.2
    q = M.unpack(p)
    q["in_scale"][:] = rng.uniform(0.5, 2.0, q["in_scale"].size)
    q["out_scale"][:] = rng.uniform(0.5, 2.0, q["out_scale"].size)
    X = M.band_features(rng.standard_normal((5, 36)), rng.standard_normal((5, 36)))
    X[2] = 0.0
    c = M.correction(p, X)
    np.testing.assert_allclose(M.correction(p, 1e-7 * X), 1e-7 * c, rtol=1e-12, atol=0)
    np.testing.assert_allclose(M.correction(p, 3e4 * X), 3e4 * c, rtol=1e-12, atol=0)
    assert np.all(c[2] == 0.0)
    # away from zero it is the plain network applied to unit-RMS input
    s = M.column_scale(p, X[:1])
    ref = M.forward(p, X[:1] / s[:, None])[2] * q["out_scale"] * s[:, None]
    np.testing.assert_allclose(c[:1], ref, rtol=1e-12)
