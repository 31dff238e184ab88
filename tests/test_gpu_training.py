"""Learned thin boundary end to end: teacher-forced data from the plain
solver, training with the HIP kernels, closed-loop calibration on one
held-out model and the reflected energy on another, against the same thin
CPML without the network."""
import pytest
import torch

pytestmark = pytest.mark.gpu


def test_trained_network_beats_plain_thin_cpml():
    """Train on small teacher-forced runs, then (a) the one-step boundary error
    left on held-out runs is below the error without the network, and (b) in
    closed loop on a held-out model the thin CPML with the network reflects
    less energy than the same thin CPML without it (same grid, same steps)."""
    from seiswave_taichi_amd import learned_abc as LA
    torch.manual_seed(0)
    net, hists = LA.train(n_runs=16, n_int=28, steps=120, iters=1500, seed=11, log=None)
    for f, h in hists.items():
        best_val = min(v for _, _, v in h)
        assert best_val < 1.0, (f, h)
    res = LA.reflection_study(net, n_int=28, steps=150, seed=99)
    assert res["learned"] < res["thin"], res
