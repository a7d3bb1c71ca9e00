"""Learned corrector guess for the triply periodic and plane Couette steppers.

No counterpart in the reference.  The reference starts every step's
corrector from the Euler predictor ``u^1 = (u^n ldt_1 + N^n) ildt_2``
(``flows/monochromatic.py:320-333``) and iterates the incremental
Crank-Nicolson correction ``u <- u + c ildt_2 (N(u) - N_prev)``
(``flows/monochromatic.py:336-355``) until the correction's norm is below
the tolerance (``__main__.py:161-191``).  The fixed point is
``u* = u^1 + c ildt_2 (N(u*) - N^n)``.

Here the first corrector pass starts from ``u^1 + c ildt_2 dN`` with
``rhs_prev = N^n + dN``, where ``dN`` estimates ``N(u*) - N^n``; the
incremental pass from there is exactly the full form
``u <- ildt_2 (ldt_1 u^n + c N(u) + (1 - c) N^n)``.  ``dN`` is the
quadratic extrapolation ``2 N^n - 3 N^{n-1} + N^{n-2}`` of the nonlinear
term's time series, plus a per-mode network's estimate of the remaining
third difference.  The iteration map, the fixed point and the stopping
rule are the reference's; only the starting point moves, so a good guess
saves whole RHS evaluations (the plain scheme spends ``2 + c`` per step,
the guessed one ``1 + passes``).

Plane Couette (``flows/plane_couette.py:507-539``): the corrector's explicit
term is the previous evaluation, so its fixed point is ``u* = S(u^n;
N(u*))`` with ``S`` the influence-matrix solve.  The guess is the
predictor's solve fed with ``N_ext = N^n + dN`` and ``rhs_prev = N_ext``,
with the same ``dN`` per point (kz, kx, y) (y itself as the network's third
coordinate).

The network (``Net``) is trained on the GPU with the library's own FP32
kernels on samples of the solver's nonlinear-term history (``collect``,
``make_dataset``, ``train``) and runs in bf16 on the matrix cores inside
the guess kernel.
"""

from __future__ import annotations

import ctypes
import math
from dataclasses import dataclass, field

import torch

from . import _lib
from .driver import Escalation
from .native import _ptr, _stream


class Net:
    """Parameters, Adam moments and packed inference operands of one network."""

    D_IN, D_OUT = 16, 6

    def __init__(self, hidden: int = 32, seed: int = 0, device=None):
        if hidden not in (32, 64):
            raise ValueError("hidden width must be 32 or 64")
        self.lib = _lib.load()
        self.H = hidden
        n, wbytes, nbias = ctypes.c_longlong(), ctypes.c_longlong(), ctypes.c_longlong()
        _lib.check(self.lib.dnsmi_mlp_sizes(hidden, ctypes.byref(n), ctypes.byref(wbytes), ctypes.byref(nbias)))
        self.nparams = n.value
        device = torch.device("cuda", torch.cuda.current_device()) if device is None else device
        self.device = device
        g = torch.Generator().manual_seed(seed)
        H, di, do = hidden, self.D_IN, self.D_OUT
        parts = [torch.randn(di, H, generator=g, dtype=torch.float64) / math.sqrt(di), torch.zeros(H),
                 torch.randn(H, H, generator=g, dtype=torch.float64) / math.sqrt(H), torch.zeros(H),
                 torch.randn(H, do, generator=g, dtype=torch.float64) * (0.1 / math.sqrt(H)), torch.zeros(do)]
        self.params = torch.cat([p.reshape(-1).float() for p in parts]).to(device)
        assert self.params.numel() == self.nparams
        self.m = torch.zeros_like(self.params)
        self.v = torch.zeros_like(self.params)
        self.grad = torch.zeros_like(self.params)
        self.loss = torch.zeros((), dtype=torch.float32, device=device)
        self.steps = 0
        self.wb = torch.zeros(wbytes.value, dtype=torch.uint8, device=device)
        self.bias = torch.zeros(nbias.value, dtype=torch.float32, device=device)
        self._part = torch.empty(0, dtype=torch.float32, device=device)

    # ------------------------------------------------------------ parameters
    def unpack(self):
        """(W1, b1, W2, b2, W3, b3) views of the FP32 parameters."""
        H, di, do = self.H, self.D_IN, self.D_OUT
        shapes = [(di, H), (H,), (H, H), (H,), (H, do), (do,)]
        out, o = [], 0
        for s in shapes:
            k = math.prod(s)
            out.append(self.params[o:o + k].view(s))
            o += k
        return out

    def set_params(self, flat: torch.Tensor) -> None:
        self.params.copy_(flat.to(self.params))

    # ------------------------------------------------------------ kernels
    def compute_grad(self, X: torch.Tensor, T: torch.Tensor, bf16: bool = True) -> torch.Tensor:
        """FP32 gradient of the mean squared error into ``self.grad``; returns the loss (device)."""
        B = X.shape[0]
        need = ((B + 31) // 32 + (B + 1023) // 1024) * (self.nparams + 1)
        if self._part.numel() < need:
            self._part = torch.empty(need, dtype=torch.float32, device=self.device)
        _lib.check(self.lib.dnsmi_mlp_grad(self.H, B, _ptr(X), _ptr(T), _ptr(self.params), int(bf16),
                                           _ptr(self._part), _ptr(self.grad), _ptr(self.loss), _stream()))
        return self.loss

    def adam(self, lr: float) -> None:
        self.steps += 1
        _lib.check(self.lib.dnsmi_mlp_adam(self.H, _ptr(self.params), _ptr(self.grad), _ptr(self.m),
                                           _ptr(self.v), self.steps, lr, _stream()))

    def pack(self) -> None:
        """Refresh the bf16 matrix-core operands from the FP32 parameters."""
        _lib.check(self.lib.dnsmi_mlp_pack(self.H, _ptr(self.params), _ptr(self.wb), _ptr(self.bias),
                                           _stream()))

    def forward(self, X: torch.Tensor) -> torch.Tensor:
        """Matrix-core inference on packed operands (call ``pack`` first)."""
        Y = torch.empty((X.shape[0], self.D_OUT), dtype=torch.float32, device=self.device)
        _lib.check(self.lib.dnsmi_mlp_forward(self.H, _ptr(X), X.shape[0], _ptr(self.wb), _ptr(self.bias),
                                              _ptr(Y), _stream()))
        return Y

    @property
    def operands(self):
        return (self.H, self.wb, self.bias)


# ---------------------------------------------------------------- data
def collect(flow, state: torch.Tensor, n_steps: int):
    """Run the plain stepper for n_steps and return the nonlinear terms
    N^0..N^n at the start of every step (plus the final state)."""
    from .driver import run

    snaps = [flow.get_rhs(state).clone()]
    res = run(flow, state.clone(), n_steps, callback=lambda i, s: snaps.append(flow.get_rhs(s).clone()))
    if not res.converged:
        raise RuntimeError("data run did not converge")
    return snaps, res.state


def _rms(t: torch.Tensor) -> float:
    return float(torch.sqrt(torch.mean(t.real**2 + t.imag**2) / 2))


@dataclass
class Scales:
    """Feature scales s_1, s_2 and target scale s_out; ``cubic``: the
    targets are residuals of cubic extrapolation (the guess then builds on
    cubic extrapolation and keeps three histories), else of quadratic."""
    s_1: float
    s_2: float
    s_out: float
    cubic: bool = False

    def astuple(self):
        return (self.s_1, self.s_2, self.s_out)


def fit_scales(snaps, cubic: bool = False) -> Scales:
    """Feature and target scales over the collected history: 1/rms(D1),
    1/rms(D2) and rms(D3) (rms(D4) for cubic targets), Dk the k-th
    difference of N."""
    n = len(snaps)
    if n < (5 if cubic else 4):
        raise ValueError("need at least four (cubic: five) nonlinear terms")
    d1 = sum(_rms(snaps[i] - snaps[i - 1]) for i in range(1, n)) / (n - 1)
    d2 = sum(_rms(snaps[i] - 2 * snaps[i - 1] + snaps[i - 2]) for i in range(2, n)) / (n - 2)
    if cubic:
        out = sum(_rms(snaps[i] - 4 * snaps[i - 1] + 6 * snaps[i - 2] - 4 * snaps[i - 3] + snaps[i - 4])
                  for i in range(4, n)) / (n - 4)
    else:
        out = sum(_rms(snaps[i] - 3 * snaps[i - 1] + 3 * snaps[i - 2] - snaps[i - 3]) for i in range(3, n)) / (n - 3)
    return Scales(1.0 / max(d1, 1e-300), 1.0 / max(d2, 1e-300), max(out, 1e-300), cubic)


def make_dataset(flow, snaps, scales: Scales, samples_per_window: int, seed: int = 0):
    """Features X (S,16) and targets T (S,6) from every window of four
    (cubic targets: five) consecutive nonlinear terms, at uniformly drawn
    modes."""
    nat = flow._native
    n_modes = math.prod(flow.params.spec_shape)
    g = torch.Generator(device=nat.device).manual_seed(seed)
    Xs, Ts = [], []
    for i in range(3 if scales.cubic else 2, len(snaps) - 1):
        modes = torch.randint(0, n_modes, (samples_per_window,), generator=g, device=nat.device,
                              dtype=torch.int64)
        X = torch.empty((samples_per_window, Net.D_IN), dtype=torch.float32, device=nat.device)
        T = torch.empty((samples_per_window, Net.D_OUT), dtype=torch.float32, device=nat.device)
        nat.mlp_features(snaps[i - 2], snaps[i - 1], snaps[i], snaps[i + 1], modes, *scales.astuple(), X, T,
                         n_m3=snaps[i - 3] if scales.cubic else None)
        Xs.append(X)
        Ts.append(T)
    return torch.cat(Xs), torch.cat(Ts)


def train(net: Net, X: torch.Tensor, T: torch.Tensor, iters: int, batch: int = 8192, lr: float = 2e-3,
          seed: int = 0, bf16: bool = True, log_every: int = 0):
    """Mini-batch Adam with a cosine learning-rate decay; returns the loss history."""
    g = torch.Generator(device=X.device).manual_seed(seed)
    hist = []
    for it in range(iters):
        idx = torch.randint(0, X.shape[0], (batch,), generator=g, device=X.device)
        net.compute_grad(X[idx].contiguous(), T[idx].contiguous(), bf16=bf16)
        net.adam(lr * 0.5 * (1 + math.cos(math.pi * it / iters)))
        if log_every and (it % log_every == 0 or it == iters - 1):
            hist.append((it, float(net.loss)))
    net.pack()
    return hist


def mean_loss(net: Net, X: torch.Tensor, T: torch.Tensor) -> float:
    """Loss of the packed (bf16 matrix-core) network on a held-out set, next
    to the loss of predicting zero (quadratic extrapolation alone)."""
    Y = net.forward(X)
    return float(torch.mean(torch.sum((Y - T) ** 2, dim=1))), float(torch.mean(torch.sum(T**2, dim=1)))


# ---------------------------------------------------------------- stepping
@dataclass
class GuessRunResult:
    state: object
    passes: list = field(default_factory=list)   # corrector passes per step (1 RHS each)
    errors: list = field(default_factory=list)
    converged: bool = True
    nan: bool = False
    escalations: int = 0
    rerecorded: int = 0

    @property
    def rhs_evals(self) -> int:
        return sum(c + 1 for c in self.passes)


MODES = {"plain": 0, "linear": 1, "learned": 2, "quadratic": 3, "cubic": 4}


def _cubic(mode: int, net, scales) -> bool:
    """Whether a periodic guess of this mode keeps three histories."""
    return mode == 4 or (mode == 2 and not hasattr(net, "guess") and scales is not None and scales.cubic)


def _startup_mode(it: int, mode: int, cubic: bool) -> int:
    """Mode of step `it`: the plain predictor, then linear, then (cubic runs)
    quadratic extrapolation until the history is long enough."""
    if it == 0:
        return 0
    if it == 1:
        return min(mode, 1)
    if it == 2 and cubic:
        return 3
    return mode


def run_guessed(flow, state, n_steps: int, mode: int = 2, net: Net | None = None,
                scales: Scales | None = None, callback=None) -> GuessRunResult:
    """The reference's time loop (``__main__.py:125-203``) with the corrector
    started from a guess: mode 1 linear, 3 quadratic, 4 cubic extrapolation
    of N (periodic flows), 2 quadratic (cubic when ``scales.cubic``) plus
    the network (for plane Couette the guess is the
    predictor's solve fed with the extrapolated N, and the network acts per
    point (kz, kx, y), or per mode (kz, kx) on whole wall-normal profiles
    when ``net`` is a ``profile_net.ProfileNet``).  Mode 0 is the plain
    predictor: the same steps as the
    plain driver, counted as passes (= iterations + 1).  The first steps of a run lack history and
    start from the plain predictor (step 0), linear extrapolation (step 1)
    and, for cubic runs, quadratic extrapolation (step 2).  Up to
    ``max_corrector_iterations + 1`` passes per step (the plain
    scheme's RHS budget); the loop stops after a step that did not
    converge, and the mixed form's escalation to FP64 passes applies as in
    ``driver.run``.  The state handed to ``callback`` is a buffer that
    later steps reuse: copy it to keep it."""
    if mode not in (0, 1, 2, 3, 4):
        raise ValueError("mode must be 0 to 4")
    if mode == 2 and (net is None or scales is None):
        raise ValueError("mode 2 needs a trained net and its scales")
    p = flow.params
    nat = flow._native
    tol, max_pass = p.corrector_tolerance, p.max_corrector_iterations + 1
    esc = Escalation(flow)
    sc = (1.0, 1.0, 1.0) if scales is None else scales.astuple()
    ops = None if net is None or hasattr(net, "guess") else net.operands
    cubic = _cubic(mode, net, scales)
    carry = not p.walls  # periodic passes run in the carry form (driver.carry_context)
    hist = [torch.zeros(nat.shape + (2,), dtype=torch.float32, device=nat.device) for _ in range(3 if cubic else 2)]
    res = GuessRunResult(state=state)
    state = state.clone()
    pred = nat.empty()
    rhs = [nat.empty(), nat.empty()]
    err_t = nat.scalar()
    last_error = 0.0
    for it in range(n_steps):
        if not last_error < tol:
            break
        m = _startup_mode(it, mode, cubic)
        if m == 2 and hasattr(net, "guess"):  # the plane Couette profile network
            net.guess(nat, state, pred, rhs[0], hist[0], hist[1], scales)
        else:
            nat.guess(state, pred, rhs[0], hist[0], hist[1], m, ops, sc, hist3=hist[2] if cubic else None,
                      carry=carry)
        if cubic:
            hist = [hist[2], hist[0], hist[1]]  # N^n went to the oldest field
        else:
            hist.reverse()  # hist[0] = N^n, hist[1] = N^{n-1}
        c, err, cur = 0, math.inf, 0
        esc.begin()
        while err > tol and c < max_pass:
            if carry:
                nat.iterate_correction_carry(pred, rhs[0], err_t)
            else:
                nat.iterate_correction(state, pred, rhs[cur], rhs[1 - cur], err_t)
                cur = 1 - cur
            err = float(err_t)
            c += 1
            if c < max_pass:
                esc.after(err)
        res.escalations += int(esc.end())
        res.rerecorded = esc.recorded
        pred, state = state, pred
        state, _ = flow.correct_velocity(state)
        res.passes.append(c)
        res.errors.append(err)
        last_error = err
        if callback is not None:
            callback(it + 1, state)
        if math.isnan(err):
            res.nan = True
            break
    res.state = state
    res.converged = not (last_error > tol) and not res.nan
    return res


class GuessedGraphStepper:
    """Guessed steps with the corrector decided on the device, one captured
    graph launch per step (two graphs: state and history buffers swap each
    step; six for cubic guesses, whose three histories rotate with period
    three).  The start-up steps (plain predictor, linear extrapolation and,
    for cubic guesses, quadratic extrapolation) run eagerly through the same
    device step with the full pass budget; the
    rest replay the graph of the requested mode with `slots` corrector slots
    (default: the full budget, max_corrector_iterations + 1).  A step that
    needs more passes than the graph holds is counted in
    ``step_stats()["unconverged"]``; a run with none is identical to the
    eager loop."""

    def __init__(self, flow, state, mode: int = 2, net: Net | None = None, scales: Scales | None = None,
                 slots: int | None = None, capacity: int = 4096):
        if mode not in (1, 2, 3, 4):
            raise ValueError("mode must be 1 to 4")
        if mode == 2 and (net is None or scales is None):
            raise ValueError("mode 2 needs a trained net and its scales")
        nat = flow._native
        self.nat, self.mode = nat, mode
        p = flow.params
        self.slots = p.max_corrector_iterations + 1 if slots is None else slots
        self.ops = None if net is None else net.operands
        self.sc = (1.0, 1.0, 1.0) if scales is None else scales.astuple()
        self.bufs = [state.detach().clone().contiguous(), nat.empty()]
        self.rhs = [nat.empty(), nat.empty()]
        self.cubic = _cubic(mode, net, scales)
        self.hist = [torch.zeros(nat.shape + (2,), dtype=torch.float32, device=nat.device)
                     for _ in range(3 if self.cubic else 2)]
        self.counts = torch.zeros(capacity, dtype=torch.int32, device=nat.device)
        self.capacity = capacity
        self.parity = 0
        self.steps = 0
        nat.step_reset()
        # the start-up steps run eagerly with the full pass budget
        full = p.max_corrector_iterations + 1
        self.period = 6 if self.cubic else 2
        start = 3 if self.cubic else 2
        for it in range(start):
            self._enqueue(it, _startup_mode(it, mode, self.cubic), full)
        self.steps = start
        self.parity = start % 2
        self.graphs = [self._capture(k) for k in range(self.period)]

    def _enqueue(self, k: int, mode: int, slots: int | None = None) -> None:
        """Step k's launches: state buffers by k mod 2; histories by k mod 2
        (two, swapped) or k mod 3 (three, rotated: step k writes N into the
        oldest, h[(2 - k) mod 3])."""
        b, h = self.bufs, self.hist
        parity = k % 2
        if self.cubic:
            r = k % 3
            h1, h2, h3 = h[(3 - r) % 3], h[(4 - r) % 3], h[(5 - r) % 3]
        else:
            h1, h2, h3 = h[parity], h[1 - parity], None
        self.nat.step_device_guess(b[parity], b[1 - parity], self.rhs[0], self.rhs[1], h1, h2, mode,
                                   self.slots if slots is None else slots, self.counts, self.ops, self.sc,
                                   hist3=h3)

    def _capture(self, k: int):
        g = torch.cuda.CUDAGraph()
        torch.cuda.synchronize()
        with torch.cuda.graph(g):
            self._enqueue(k, self.mode)
        return g

    def step(self, n: int = 1) -> None:
        if self.steps + n > self.capacity:
            raise ValueError("pass-count capacity exceeded")
        for _ in range(n):
            self.graphs[self.steps % self.period].replay()
            self.steps += 1
            self.parity = self.steps % 2

    @property
    def state(self):
        return self.bufs[self.parity]

    def passes(self):
        return self.counts[: self.steps].tolist()

    def step_stats(self):
        return self.nat.step_stats()
