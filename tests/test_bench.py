"""The benchmark's byte accounting and its command-line contract, on the host.

The algorithmic bytes of a step (two-sweep schedule, f32) are 120 B per
updated cell plus 48 B per (updated cell, PML direction) pair.  Here the
counts are checked against a brute-force count over every cell of small
grids, and against the figures of the 200^3 case with a 20-cell CPML."""
import json
import os
import subprocess
import sys

import numpy as np
import pytest

from tests.conftest import ROOT

sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _brute(shape, s, slabs):
    """N_u and S_u by enumerating the updated cells [s, n-s-1) of each axis
    and testing each axis index against its slab bounds."""
    idx = [np.arange(s, n - s - 1) for n in shape]
    grid = np.meshgrid(*idx, indexing="ij")
    n_u = grid[0].size
    s_u = sum(int(((g < lo) | (g >= hi)).sum()) for g, (lo, hi) in zip(grid, slabs))
    return n_u, s_u


@pytest.mark.parametrize("shape,s,slabs", [
    ((30, 26, 22), 2, [(6, 24), (6, 20), (6, 16)]),
    ((17, 40, 9), 1, [(3, 14), (0, 40), (2, 7)]),    # no slab on y
    ((24, 24, 24), 4, [(24, 24), (9, 15), (9, 15)]),  # x all PML
])
def test_algorithmic_counts_match_a_brute_force_count(shape, s, slabs):
    n_u, s_u, vel, stress = bench.algorithmic_bytes(shape, s, slabs)
    assert (n_u, s_u) == _brute(shape, s, slabs)
    assert vel == 52 * n_u + 24 * s_u and stress == 68 * n_u + 24 * s_u
    n_u2, s_u2, vel2, stress2 = bench.algorithmic_bytes(shape, s, slabs, spml=True)
    assert (n_u2, s_u2) == (n_u, s_u)
    assert vel2 == 52 * n_u + 32 * s_u and stress2 == 68 * n_u + 32 * s_u


def test_200_cube_step_bytes():
    """200^3, s = 2, 10 m, 20-cell CPML (200 m): the 195^3 updated cells.
    The PML coordinate of index i is (i - 2) * 10 m (+5 m at half points);
    the lo layer holds x < 200 m, i.e. i <= 21, and the hi layer x > 1750 m
    (xmax = 1950 m), which half points reach from i = 177.  So the memory
    slabs are [0, 22) and [177, 200): 20 + 20 of the 195 updated indices per
    axis, S_u = 3 * 40 * 195^2, and the step moves 120 N_u + 48 S_u bytes."""
    from seiswave_taichi_amd.pml import axis_pml
    w = bench.workload("tensor_200", 10)
    p = w["pml"]
    ax = axis_pml(200, 2, 10.0, w["dt"], p["pml_x_thick"], p["vp_max"], p["Rcoef"], p["alpha_max_pml"],
                  p["kmax_pml"], p["theta"], True, True)
    assert (ax.lo, ax.hi) == (22, 177)
    n_u, s_u, vel, stress = bench.algorithmic_bytes((200, 200, 200), 2, [(ax.lo, ax.hi)] * 3)
    assert n_u == 195 ** 3 == 7_414_875
    assert s_u == 3 * 40 * 195 ** 2 == 4_563_000
    assert vel + stress == 120 * n_u + 48 * s_u == 1_108_809_000


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 2
    line = json.loads(r.stdout.strip().splitlines()[-1])
    assert line["n_gpus"] == 2 and "one GPU" in line["error"]
