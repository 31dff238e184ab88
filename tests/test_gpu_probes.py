"""The measurement probes the bench reports its ceilings with, against their
closed forms: the streaming probe bit for bit, and the matrix-core probe
against the f32 16x16x4 MFMA operand layout the band correction relies on
(lane li + 16 g supplies A[li][g] and B[g][li] and receives D[4 g + r][li])."""
import ctypes as C

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("nin,nout", [(10, 3), (11, 6), (16, 12), (20, 21), (1, 1)])
def test_stream_probe_closed_form(nin, nout):
    from seiswave_taichi_amd import _lib as L
    lib = L.load()
    n = 4096 + 64
    rng = np.random.default_rng(nin * 100 + nout)
    host_in = [rng.standard_normal(n).astype(np.float32) for _ in range(nin)]
    dev_in = [torch.from_numpy(a).cuda() for a in host_in]
    dev_out = [torch.empty(n, dtype=torch.float32, device="cuda") for _ in range(nout)]
    pin = (L._fp * nin)(*[L.fptr(t) for t in dev_in])
    pout = (L._fp * nout)(*[L.fptr(t) for t in dev_out])
    st = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(lib.sw_stream_probe(pin, nin, pout, nout, n, 7, st), "stream probe")
    torch.cuda.synchronize()
    acc = np.zeros(n, np.float32)
    for a in host_in:  # the kernel's order: sum over the inputs from 0, then the scale
        acc = acc + a
    for o, t in enumerate(dev_out):
        np.testing.assert_array_equal(t.cpu().numpy(), acc * np.float32(o + 1))


def test_stream_probe_refuses_other_mixes():
    from seiswave_taichi_amd import _lib as L
    lib = L.load()
    bufs = [torch.zeros(64, device="cuda") for _ in range(5)]
    pin = (L._fp * 3)(*[L.fptr(t) for t in bufs[:3]])
    pout = (L._fp * 2)(*[L.fptr(t) for t in bufs[3:]])
    rc = lib.sw_stream_probe(pin, 3, pout, 2, 64, 1, C.c_void_p(torch.cuda.current_stream().cuda_stream))
    assert rc == -1  # SW_EINVAL


def test_mfma_probe_operand_layout():
    from seiswave_taichi_amd import _lib as L
    lib = L.load()
    nblock, steps = 3, 8
    out = torch.empty(nblock * 256, dtype=torch.float32, device="cuda")
    L.check(lib.sw_mfma_probe(L.fptr(out), nblock, steps, C.c_void_p(torch.cuda.current_stream().cuda_stream)),
            "matrix-core probe")
    torch.cuda.synchronize()
    got = out.cpu().numpy().reshape(nblock, 4, 64)
    lane = np.arange(64, dtype=np.float32)
    a = (np.float32(1.0) + np.float32(0.01) * lane).astype(np.float64)   # lane l supplies A and
    b = (np.float32(1.0) - np.float32(0.005) * lane).astype(np.float64)  # B operands a[l], b[l]
    A = np.array([[a[i + 16 * g] for g in range(4)] for i in range(16)])  # A[i][g] from lane i + 16 g
    B = np.array([[b[j + 16 * g] for j in range(16)] for g in range(4)])  # B[g][j] from lane j + 16 g
    D = A @ B
    want = np.array([4 * steps * sum(D[4 * (ln // 16) + r, ln % 16] for r in range(4)) for ln in range(64)])
    # a transposed or shifted layout moves these sums by per cent, the f32 accumulation by ~1e-6
    assert np.abs(want - want[::-1]).max() > 1e-2 * np.abs(want).max()
    for blk in range(nblock):
        for w in range(4):
            np.testing.assert_allclose(got[blk, w], want, rtol=1e-5)
