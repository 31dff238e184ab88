"""The allocation-free MLP backward (sw_mlp_loss_backward_noalloc) against
sw_mlp_loss_backward with a caller-owned work buffer: the same kernels on the
same inputs, so gradients and per-row losses are equal bit for bit, on two
calls in a row (the second with fewer rows, so it runs over the scratch the
first call left behind)."""
import ctypes as C

import numpy as np
import pytest
import torch

from seiswave_taichi_amd import _lib as L

pytestmark = pytest.mark.gpu

NIN, NH, NOUT = 72, 64, 36


def _call_pair(lib, params, rng, B):
    dev = params.device
    X = torch.from_numpy(rng.standard_normal((B, NIN)).astype(np.float32)).to(dev)
    T = torch.from_numpy(rng.standard_normal((B, NOUT)).astype(np.float32)).to(dev)
    H1, H2 = torch.empty((B, NH), device=dev), torch.empty((B, NH), device=dev)
    Y = torch.empty((B, NOUT), device=dev)
    stream = C.c_void_p(torch.cuda.current_stream().cuda_stream)
    L.check(lib.sw_mlp_forward(L.fptr(params), L.fptr(X), B, L.fptr(H1), L.fptr(H2), L.fptr(Y), stream), "forward")
    ngrad = lib.sw_mlp_param_count() - NIN - NOUT
    g_ref, g_new = torch.full((ngrad,), np.nan, device=dev), torch.full((ngrad,), np.nan, device=dev)
    r_ref, r_new = torch.full((B,), np.nan, device=dev), torch.full((B,), np.nan, device=dev)
    work = torch.empty(B * (NOUT + 2 * NH), device=dev)
    args = [L.fptr(t) for t in (params, X, H1, H2, Y, T)]
    L.check(lib.sw_mlp_loss_backward(*args, B, L.fptr(g_ref), L.fptr(work), L.fptr(r_ref), stream), "backward")
    L.check(lib.sw_mlp_loss_backward_noalloc(*args, B, L.fptr(g_new), L.fptr(r_new), stream), "backward noalloc")
    torch.cuda.synchronize()
    return g_ref.cpu().numpy(), g_new.cpu().numpy(), r_ref.cpu().numpy(), r_new.cpu().numpy()


def test_noalloc_backward_equals_work_buffer_backward():
    lib = L.load()
    torch.cuda.set_device(0)
    L.check(lib.sw_device_init(0), "device init")
    rng = np.random.default_rng(3)
    n = lib.sw_mlp_param_count()
    p = rng.standard_normal(n).astype(np.float32) * 0.1
    p[n - NIN - NOUT:] = 1.0 + 0.1 * rng.random(NIN + NOUT).astype(np.float32)  # input / output scales
    params = torch.from_numpy(p).cuda()
    for B in (3000, 1000):
        print(f"noalloc backward, B={B}", flush=True)
        g_ref, g_new, r_ref, r_new = _call_pair(lib, params, rng, B)
        assert np.all(np.isfinite(g_ref)) and np.abs(g_ref).max() > 0
        np.testing.assert_array_equal(g_new, g_ref)
        np.testing.assert_array_equal(r_new, r_ref)


def test_noalloc_backward_refuses_batches_beyond_its_scratch():
    lib = L.load()
    fake = C.cast(C.c_void_p(0x1000), L._fp)
    B = L.SW_MLP_NOALLOC_MAX_ROWS + 1
    rc = lib.sw_mlp_loss_backward_noalloc(*([fake] * 6), B, fake, fake, None)
    assert rc == -1  # SW_EINVAL
    assert "allocation-free scratch" in L.last_error()
