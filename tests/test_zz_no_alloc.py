"""bc_closure_grad_noalloc (partial sums in the library's fixed device array)
against bc_closure_grad (caller workspace) through the C ABI: bitwise equal
gradients and losses on two calls in a row, the second with a smaller batch,
and a clear error above the array's batch bound."""

import ctypes

import numpy as np
import pytest

from braincell_amd import _native
from braincell_amd.closure import NPARAM, ClosureNet, _TrainDesc

pytestmark = pytest.mark.gpu

MAX_BATCH = 32768  # BC_CLOSURE_NOALLOC_MAX_BATCH


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def test_noalloc_grad_equals_workspace_grad(cuda):
    import torch

    L = _native.lib()
    g = torch.Generator(device="cpu").manual_seed(5)
    n = 50_000
    X = torch.randn(n, 8, generator=g).cuda()
    Dn = torch.randn(n, 5, generator=g).cuda()
    Yn = (0.1 * torch.randn(n, 5, generator=g)).cuda()
    P = torch.as_tensor(ClosureNet.init_params(3, out_gain=0.3), dtype=torch.float32).cuda()
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    for batch in (MAX_BATCH - 77, 3000):  # two calls in a row, the second smaller
        print(f"noalloc grad: batch {batch}", flush=True)
        idx = torch.randperm(n, generator=g)[:batch].cuda()
        nblk = (batch + 127) // 128
        partial = torch.empty((nblk, NPARAM), device="cuda")
        loss_part = torch.empty(nblk, device="cuda")
        grad_ref, loss_ref = torch.empty(NPARAM, device="cuda"), torch.empty(1, device="cuda")
        d = _TrainDesc(batch, nblk, _ptr(X), _ptr(Dn), _ptr(Yn), _ptr(idx), _ptr(P), _ptr(partial), _ptr(loss_part),
                       _ptr(grad_ref), _ptr(loss_ref))
        _native.check(L.bc_closure_grad(ctypes.byref(d), stream))
        grad, loss = torch.full((NPARAM,), float("nan"), device="cuda"), torch.full((1,), float("nan"), device="cuda")
        d2 = _TrainDesc(batch, 0, _ptr(X), _ptr(Dn), _ptr(Yn), _ptr(idx), _ptr(P), None, None, _ptr(grad), _ptr(loss))
        _native.check(L.bc_closure_grad_noalloc(ctypes.byref(d2), stream))
        torch.cuda.synchronize()
        assert torch.isfinite(grad_ref).all() and float(grad_ref.abs().max()) > 0
        assert torch.equal(grad, grad_ref), float((grad - grad_ref).abs().max())
        assert torch.equal(loss, loss_ref)


def test_noalloc_grad_rejects_batches_beyond_the_array(cuda):
    L = _native.lib()
    d = _TrainDesc(MAX_BATCH + 1, 0, *([ctypes.c_void_p(16)] * 5), None, None, ctypes.c_void_p(16),
                   ctypes.c_void_p(16))
    assert L.bc_closure_grad_noalloc(ctypes.byref(d), None) == 1
    assert b"exceeds" in L.bc_last_error()
