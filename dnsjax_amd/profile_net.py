"""Per-(kz, kx) profile network for the plane Couette corrector guess.

No counterpart in the reference.  The wall-bounded corrector's fixed point
is ``u* = S(u^n; N(u*))`` (``flows/plane_couette.py:507-539``), so the best
starting iterate is the predictor's solve fed with the nonlinear term of
the converged state, which is the next step's ``N^{n+1}``.  The guess feeds
the solve with quadratic extrapolation of the nonlinear term plus this
network's estimate of the extrapolation's residual, mode by mode: the input
is the whole wall-normal profile of the mode's first and second differences
(12 Ny values plus three mode features), the output the residual's profile
(6 Ny values), with two hidden layers of width H (256 by default).  Training
pairs (predictor state's history -> converged corrector state's N) come
from the solver's own steps; training and inference run every layer,
forward and backward, as bf16 matrix-core GEMMs with FP32 accumulation
(``csrc/profile_net.hip``).  The corrector then iterates to its tolerance
as before, so only its starting point moves.
"""

from __future__ import annotations

import ctypes
import math

import torch

from . import _lib
from .learned import Scales, _rms
from .native import _ptr, _stream


def profile_din(ny: int) -> int:
    """Input width: 12 Ny profile values and three mode features, padded to a multiple of 32."""
    return -(-(12 * ny + 3) // 32) * 32


class ProfileNet:
    """FP32 parameters, Adam moments and the bf16 weight copies of one
    profile network for a plane Couette flow of ``ny`` points."""

    def __init__(self, ny: int, hidden: int = 256, seed: int = 0, device=None, shape=None):
        """``shape`` = (din, dout) overrides the widths a flow of ``ny``
        points implies (a network used outside the guess)."""
        if hidden % 64:
            raise ValueError("hidden width must be a multiple of 64")
        self.lib = _lib.load()
        self.ny, self.H = ny, hidden
        self.din, self.dout = (profile_din(ny), 6 * ny) if shape is None else shape
        n = self.lib.dnsmi_pnet_params(self.din, hidden, self.dout)
        if n < 0:
            raise ValueError("unsupported profile network shape")
        self.nparams = n
        device = torch.device("cuda", torch.cuda.current_device()) if device is None else device
        self.device = device
        g = torch.Generator().manual_seed(seed)
        H, di, do = hidden, self.din, self.dout
        parts = [torch.randn(di, H, generator=g, dtype=torch.float64) / math.sqrt(di), torch.zeros(H),
                 torch.randn(H, H, generator=g, dtype=torch.float64) / math.sqrt(H), torch.zeros(H),
                 torch.randn(H, do, generator=g, dtype=torch.float64) * (0.1 / math.sqrt(H)), torch.zeros(do)]
        self.params = torch.cat([p.reshape(-1).float() for p in parts]).to(device)
        assert self.params.numel() == n
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.wb = torch.zeros(n, dtype=torch.int16, device=device)
        self.loss = torch.zeros((), dtype=torch.float32, device=device)
        self.steps = 0
        self._work = {}

    def _buf(self, name, numel, dtype):
        t = self._work.get(name)
        if t is None or t.numel() < numel:
            t = torch.empty(numel, dtype=dtype, device=self.device)
            self._work[name] = t
        return t

    def unpack(self):
        """(W1, b1, W2, b2, W3, b3) views of the FP32 parameters."""
        H, di, do = self.H, self.din, self.dout
        shapes = [(di, H), (H,), (H, H), (H,), (H, do), (do,)]
        out, o = [], 0
        for s in shapes:
            k = math.prod(s)
            out.append(self.params[o:o + k].view(s))
            o += k
        return out

    def compute_grad(self, X: torch.Tensor, T: torch.Tensor) -> torch.Tensor:
        """Gradient of the mean squared error into ``self.grad`` (bf16
        matrix-core GEMMs, FP32 accumulation); returns the loss (device)."""
        B = X.shape[0]
        wf = self._buf("wf", B * (2 * self.H + self.dout), torch.float32)
        wh = self._buf("wh", B * (3 * self.H + self.dout), torch.int16)
        part = self._buf("part", (B * self.dout + 255) // 256, torch.float32)
        _lib.check(self.lib.dnsmi_pnet_grad(self.din, self.H, self.dout, B, _ptr(X), _ptr(T), _ptr(self.params),
                                            _ptr(self.wb), _ptr(wf), _ptr(wh), _ptr(part), _ptr(self.grad),
                                            _ptr(self.loss), _stream()))
        return self.loss

    def adam(self, lr: float) -> None:
        self.steps += 1
        _lib.check(self.lib.dnsmi_pnet_adam(self.nparams, _ptr(self.params), _ptr(self.grad), _ptr(self.m),
                                            _ptr(self.v), self.steps, lr, _stream()))

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        B = X.shape[0]
        wf = self._buf("wf", B * 2 * self.H, torch.float32)
        wh = self._buf("wh", B * 2 * self.H, torch.int16)
        Y = torch.empty((B, self.dout), dtype=torch.float32, device=self.device)
        _lib.check(self.lib.dnsmi_pnet_forward(self.din, self.H, self.dout, B, _ptr(X), _ptr(self.params),
                                               _ptr(self.wb), _ptr(wf), _ptr(wh), _ptr(Y), _stream()))
        return Y

    def guess(self, nat, state, pred, rhs_prev, hist1, hist2, scales: Scales) -> None:
        """The plane Couette guess of ``dnsmi_guess_profile`` (hist2 <- N^n)."""
        M = nat.shape[1] * nat.shape[2]
        X = self._buf("gX", M * self.din, torch.int16)
        wf = self._buf("gwf", M * 2 * self.H, torch.float32)
        wh = self._buf("gwh", M * 2 * self.H, torch.int16)
        Y = self._buf("gY", M * self.dout, torch.float32)
        s_1, s_2, s_out = scales.astuple()
        nat._chk(self.lib.dnsmi_guess_profile(nat.handle, _ptr(state), _ptr(pred), _ptr(rhs_prev), _ptr(hist1),
                                              _ptr(hist2), self.din, self.H, self.dout, _ptr(self.params),
                                              _ptr(self.wb), _ptr(X), _ptr(wf), _ptr(wh), _ptr(Y), s_1, s_2,
                                              s_out, _stream()))


def make_dataset(flow, snaps, scales: Scales, n_pairs: int, din: int, seed: int = 0):
    """``n_pairs`` training pairs (X bf16 (n, din), T (n, 6 Ny)) drawn
    uniformly over the modes of every window of four consecutive nonlinear
    terms of the history (the fourth is the converged corrector's N)."""
    nat = flow._native
    M = nat.shape[1] * nat.shape[2]
    ny = nat.shape[3]
    windows = list(range(2, len(snaps) - 1))
    if not windows:
        raise ValueError("need at least four nonlinear terms")
    g = torch.Generator(device=nat.device).manual_seed(seed)
    per = -(-n_pairs // len(windows))
    Xs, Ts = [], []
    for i in windows:
        modes = torch.randint(0, M, (per,), generator=g, device=nat.device, dtype=torch.int64)
        X = torch.empty((per, din), dtype=torch.int16, device=nat.device)
        T = torch.empty((per, 6 * ny), dtype=torch.float32, device=nat.device)
        s_1, s_2, s_out = scales.astuple()
        nat._chk(nat.lib.dnsmi_pnet_features(nat.handle, _ptr(snaps[i - 2]), _ptr(snaps[i - 1]), _ptr(snaps[i]),
                                             _ptr(snaps[i + 1]), _ptr(modes), per, din, s_1, s_2, s_out,
                                             _ptr(X), _ptr(T), _stream()))
        Xs.append(X)
        Ts.append(T)
    return torch.cat(Xs)[:n_pairs], torch.cat(Ts)[:n_pairs]


def train(net: ProfileNet, X, T, iters: int, batch: int = 2048, lr: float = 1e-3, seed: int = 0,
          log_every: int = 0):
    """Mini-batch Adam with a cosine learning-rate decay; returns the loss history."""
    g = torch.Generator(device=X.device).manual_seed(seed)
    hist = []
    for it in range(iters):
        idx = torch.randint(0, X.shape[0], (batch,), generator=g, device=X.device)
        net.compute_grad(X[idx].contiguous(), T[idx].contiguous())
        net.adam(lr * 0.5 * (1 + math.cos(math.pi * it / iters)))
        if log_every and (it % log_every == 0 or it == iters - 1):
            hist.append((it, float(net.loss)))
    return hist


def mean_loss(net: ProfileNet, X, T):
    """Held-out loss of the network, next to the loss of predicting zero
    (quadratic extrapolation alone)."""
    Y = net.forward(X)
    return float(torch.mean((Y - T) ** 2)), float(torch.mean(T**2))


def fit_scales(snaps) -> Scales:
    """1 / rms(D1), 1 / rms(D2) and rms(D3) over the history."""
    n = len(snaps)
    d1 = sum(_rms(snaps[i] - snaps[i - 1]) for i in range(1, n)) / (n - 1)
    d2 = sum(_rms(snaps[i] - 2 * snaps[i - 1] + snaps[i - 2]) for i in range(2, n)) / (n - 2)
    d3 = sum(_rms(snaps[i] - 3 * snaps[i - 1] + 3 * snaps[i - 2] - snaps[i - 3]) for i in range(3, n)) / (n - 3)
    return Scales(1.0 / max(d1, 1e-300), 1.0 / max(d2, 1e-300), max(d3, 1e-300))
