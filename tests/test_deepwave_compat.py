"""The deepwave-style front end: parameter conversions (CPU) and argument
checks that need no GPU."""
import numpy as np
import pytest
import torch

from seiswave_taichi_amd import deepwave_compat as D


def test_conversions_round_trip_and_match_the_definitions():
    rng = np.random.default_rng(0)
    vp = torch.from_numpy(rng.uniform(2000, 5000, (7, 9)))
    vs = vp / torch.from_numpy(rng.uniform(1.6, 2.2, (7, 9)))
    rho = torch.from_numpy(rng.uniform(1800, 2800, (7, 9)))
    lamb, mu, b = D.vpvsrho_to_lambmubuoyancy(vp, vs, rho)
    torch.testing.assert_close(mu, rho * vs ** 2)
    torch.testing.assert_close(lamb, rho * (vp ** 2 - 2 * vs ** 2))
    torch.testing.assert_close(b, 1 / rho)
    vp2, vs2, rho2 = D.lambmubuoyancy_to_vpvsrho(lamb, mu, b)
    for a, c in ((vp2, vp), (vs2, vs), (rho2, rho)):
        torch.testing.assert_close(a, c, rtol=1e-12, atol=0)
    # a fluid: vs = 0, lamb = rho vp^2
    lf, mf, bf = D.vpvsrho_to_lambmubuoyancy(torch.tensor([1500.0]), torch.tensor([0.0]), torch.tensor([1000.0]))
    assert float(mf) == 0 and float(lf) == pytest.approx(1000.0 * 1500.0 ** 2)
    with pytest.raises(ValueError):
        D.lambmubuoyancy_to_vpvsrho(lamb, mu, -b)


def test_argument_checks():
    m = torch.ones(20, 20)
    with pytest.raises(NotImplementedError):
        D.elastic(m, m, m, 1.0, 1e-4, receiver_locations_p=torch.zeros(1, 1, 2))
    with pytest.raises(ValueError):
        D.elastic(m, m, m, 1.0, 1e-4, accuracy=3)
