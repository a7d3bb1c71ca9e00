"""The no-allocation gradient entry point (dnsmi_mlp_grad_static) against
dnsmi_mlp_grad with a caller-provided workspace: identical gradients and
loss on two calls in a row, the second with fewer samples (the shared
device scratch keeps the first call's contents), and a clear error above
its sample limit."""

import ctypes

import pytest
import torch

from dnsjax_amd import _lib
from dnsjax_amd.native import _ptr

MAX_SAMPLES = 8192


def _nparams(H):
    n = ctypes.c_longlong()
    _lib.check(_lib.load().dnsmi_mlp_sizes(H, ctypes.byref(n), None, None))
    return n.value


def _inputs(H, B, seed):
    g = torch.Generator().manual_seed(seed)
    X = torch.randn(B, 16, generator=g).cuda()
    T = torch.randn(B, 6, generator=g).cuda()
    prm = (0.3 * torch.randn(_nparams(H), generator=g)).cuda()
    return X, T, prm


def _grad_workspace(H, B, X, T, prm, bf16):
    lib = _lib.load()
    P = _nparams(H)
    part = torch.full((((B + 31) // 32 + (B + 1023) // 1024) * (P + 1),), float("nan"), device="cuda")
    grad = torch.zeros(P, device="cuda")
    loss = torch.zeros((), device="cuda")
    _lib.check(lib.dnsmi_mlp_grad(H, B, _ptr(X), _ptr(T), _ptr(prm), bf16, _ptr(part), _ptr(grad),
                                  _ptr(loss), None))
    return grad, loss


def _grad_static(H, B, X, T, prm, bf16):
    lib = _lib.load()
    grad = torch.zeros(_nparams(H), device="cuda")
    loss = torch.zeros((), device="cuda")
    _lib.check(lib.dnsmi_mlp_grad_static(H, B, _ptr(X), _ptr(T), _ptr(prm), bf16, _ptr(grad),
                                         _ptr(loss), None))
    return grad, loss


@pytest.mark.gpu
@pytest.mark.parametrize("H", [32, 64])
def test_static_scratch_gradient_matches_workspace_form(H):
    for bf16 in (1, 0):
        for B, seed in ((MAX_SAMPLES, 1), (1000 + H, 2)):  # the second call is smaller
            X, T, prm = _inputs(H, B, seed)
            g_ref, l_ref = _grad_workspace(H, B, X, T, prm, bf16)
            g, l = _grad_static(H, B, X, T, prm, bf16)
            torch.cuda.synchronize()
            assert torch.isfinite(g).all()
            assert torch.equal(g, g_ref), (H, B, bf16, float((g - g_ref).abs().max()))
            assert torch.equal(l, l_ref), (H, B, bf16, float(l), float(l_ref))


@pytest.mark.gpu
def test_static_scratch_rejects_too_many_samples():
    lib = _lib.load()
    H, B = 32, MAX_SAMPLES + 1
    X, T, prm = _inputs(H, 64, 3)  # never read: the size check comes first
    grad = torch.zeros(_nparams(H), device="cuda")
    rc = lib.dnsmi_mlp_grad_static(H, B, _ptr(X), _ptr(T), _ptr(prm), 1, _ptr(grad), None, None)
    assert rc == 1
    assert b"8192" in lib.dnsmi_last_error(None)
