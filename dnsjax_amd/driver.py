"""The time loop around the hot path, with the reference driver's semantics.

``run`` follows ``__main__.py:125-203``: per step one
``predict_and_correct``, then ``iterate_correction`` (periodic native flows:
their carry form, the same iteration written in full) while
``error > corrector_tolerance`` and fewer than ``max_corrector_iterations``
extra iterations ran, then ``correct_velocity``; the loop stops after a
step whose corrector did not converge.  A NaN error is reported (the
reference stops silently on it).  The error is read on the host once per
corrector pass, as in the reference.

The mixed form evaluates each corrector pass as an FP32 increment over the
step's first evaluation, so its passes converge as FP64 ones do.  If a pass
still stops converging (its error, at or below
``params.escalate_cap_value``, is not below ``params.escalate_ratio``
times the error two passes earlier), the step's remaining passes run in
FP64 (``RunResult.escalations`` counts such steps): every step ends at
error <= ``corrector_tolerance``, as in the reference.  The rounding of an
increment scales with its size; a step whose predict-and-correct error
exceeds ``params.recenter_far_value`` (large first corrections) records a
new reference once its passes have closed in to ``recenter_near_value``,
so its later increments are small again (``RunResult.rerecorded`` counts
such steps).  The device-decided step (``dnsmi_step_device``) applies the
same rules.
"""

from __future__ import annotations

import math
from dataclasses import dataclass, field


@dataclass
class RunResult:
    state: object
    counts: list = field(default_factory=list)
    errors: list = field(default_factory=list)
    converged: bool = True
    nan: bool = False
    escalations: int = 0
    rerecorded: int = 0

    @property
    def rhs_evals(self) -> int:  # __main__.py:199-201
        return sum(c + 2 for c in self.counts)


class Escalation:
    """The mixed form's switch to FP64 passes within one step (see the
    module docstring).  ``after(c, err)`` is called after pass c with its
    error and switches the flow to FP64 for the rest of the step when the
    passes stopped converging; ``end()`` switches back."""

    def __init__(self, flow):
        p = flow.params
        self.flow = flow
        self.ratio = p.escalate_ratio if p.precision == "mixed" else None
        self.cap = p.escalate_cap_value
        self.tol = p.corrector_tolerance
        split = p.precision == "mixed" and not p.plain_mixed
        self.far = p.recenter_far_value if split else math.inf
        self.near = p.recenter_near_value
        self.on = False
        self.prev = self.prev2 = math.inf
        self.pending = False
        self.recorded = 0

    def begin(self, err0: float | None = None) -> None:
        """Start of a step's passes; err0 = its predict-and-correct error,
        whose evaluation recorded the split reference (None: a step started
        from a guess, whose first pass records it)."""
        self.prev, self.prev2 = (math.inf if err0 is None else err0), math.inf
        self.pending = None if err0 is None else err0 > self.far

    def after(self, err: float) -> None:
        if self.pending is None:  # the pass that recorded the reference
            self.pending = err > self.far
        if (self.ratio is not None and not self.on and self.tol < err <= self.cap
                and err > self.ratio * self.prev2):
            self.on = True
            self.flow.escalate(True)
        if self.pending and not self.on and self.tol < err <= self.near:
            self.pending = False
            self.recorded += 1
            self.flow._native.split_rerecord()  # the next pass records
        self.prev2, self.prev = self.prev, err

    def end(self) -> bool:
        was = self.on
        if self.on:
            self.flow.escalate(False)
            self.on = False
        return was


def carry_context(flow):
    """The native context of a periodic flow, whose passes run in the carry
    form (``dnsmi_iterate_correction_carry``: each pass reads B =
    ildt_2 (ldt_1 u^n + (1 - c) N^n) and the iterate, and writes the
    iterate), or None (plane Couette, whose corrector reads the previous
    evaluation by design, and flows without a native context)."""
    nat = getattr(flow, "_native", None)
    if nat is None or flow.params.walls or not hasattr(nat, "iterate_correction_carry"):
        return None
    return nat


def run(flow, state, n_steps: int, callback=None) -> RunResult:
    p = flow.params
    tol, max_it = p.corrector_tolerance, p.max_corrector_iterations
    esc = Escalation(flow)
    res = RunResult(state=state)
    last_error = 0.0
    nat = carry_context(flow)
    if nat is not None:
        carry, err_t = nat.empty(), nat.scalar()
    for it in range(n_steps):
        if not last_error < tol:
            break
        state_prev = state
        if nat is not None:
            state = nat.empty()
            nat.predict_and_correct_carry(state_prev, state, carry, err_t)
            error = err_t
        else:
            state, rhs, error = flow.predict_and_correct(state_prev)
        err = float(error)
        esc.begin(err)
        c = 0
        while err > tol and c < max_it:
            if nat is not None:
                nat.iterate_correction_carry(state, carry, err_t)
                error = err_t
            else:
                state, rhs, error = flow.iterate_correction(state_prev, state, rhs)
            err = float(error)
            c += 1
            if c < max_it:
                esc.after(err)
        res.escalations += int(esc.end())
        res.rerecorded = esc.recorded
        state, _ = flow.correct_velocity(state)
        res.counts.append(c)
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


class StatsRecorder:
    """The reference driver's energy record and its ``stats.npz``
    (``__main__.py:107-120``, ``137-158``, ``227-242``): E' of the initial
    state at t0; at the top of every loop iteration ``it`` with
    ``it % it_stats == 0`` (so never for the state after the last step); and
    of the final state once the loop ends, when more than one step ran.  The
    file holds ``ts`` (time, accumulated by repeated ``t += dt``) and ``Eps``.

    Pass the recorder as ``run``'s callback, then call ``finish``::

        rec = StatsRecorder(flow, state, it_stats=10)
        res = run(flow, state, n, callback=rec)
        rec.finish(res.state, len(res.counts))
        rec.save("stats.npz")
    """

    def __init__(self, flow, state, it_stats: int | None = None, t0: float = 0.0, it0: int = 0):
        self.flow, self.it_stats, self.it0 = flow, it_stats, it0
        self.dt = flow.params.dt
        self._t = [t0]  # t after k steps, accumulated as the reference does
        self.ts = [t0]
        self.Eps = [self._energy(state)]
        self._last_it = None  # iteration of the last in-loop record
        self.complete = False

    def _energy(self, state) -> float:
        return float(self.flow.get_stats(state)["E'"])

    def _time(self, k: int) -> float:
        while len(self._t) <= k:
            self._t.append(self._t[-1] + self.dt)
        return self._t[k]

    def __call__(self, k: int, state) -> None:
        it = self.it0 + k
        if self.it_stats is not None and it % self.it_stats == 0 and it > self.it0:
            self.ts.append(self._time(k))
            self.Eps.append(self._energy(state))
            self._last_it = it

    def finish(self, state, k: int) -> None:
        """Record the final state after k steps (the loop's exit)."""
        it = self.it0 + k
        if it <= self.it0 + 1:
            return  # the reference writes nothing for fewer than two steps
        if self._last_it == it:  # the loop never reaches the top of iteration it
            self.ts.pop()
            self.Eps.pop()
        self.ts.append(self._time(k))
        self.Eps.append(self._energy(state))
        self.complete = True

    def save(self, path: str = "stats.npz") -> None:
        import numpy as np

        if not self.complete:
            raise ValueError("finish() a run of at least two steps first")
        np.savez(path, ts=np.array(self.ts), Eps=np.array(self.Eps))

