"""Oracle: the stepper factory and the driver's corrector loop, FP64.

Test infrastructure only (see ``oracle/__init__.py``).

``make_stepper`` restates ``timestep.py:23-122``; ``run`` restates the
predictor/corrector loop semantics of ``__main__.py:125-203`` (the host
loop that calls the hot path), returning per-step iteration counts and
errors so that the GPU path's counts can be compared one-to-one.

Pinned by: the decaying-box Beltrami mode, whose exact per-step factor
(1/dt - (1-c) k^2/Re) / (1/dt + c k^2/Re) and zero corrector count are
fixed by the scheme (tests/test_oracle_periodic.py).
"""

from __future__ import annotations

import math


def make_stepper(get_rhs_fn, predict_fn, correct_fn, norm_fn):
    """``timestep.py:23-122``."""

    def predict_and_correct(state, *args):  # timestep.py:70-94
        rhs_prev = get_rhs_fn(state, *args)
        prediction = predict_fn(state, rhs_prev, *args)
        rhs_next = get_rhs_fn(prediction, *args)
        prediction, correction = correct_fn(state, prediction, rhs_prev, rhs_next, *args)
        return prediction, rhs_next, norm_fn(correction, *args)

    def iterate_correction(state_prev, prediction, rhs_prev, *args):  # timestep.py:98-120
        rhs_next = get_rhs_fn(prediction, *args)
        prediction, correction = correct_fn(state_prev, prediction, rhs_prev, rhs_next, *args)
        return prediction, rhs_next, norm_fn(correction, *args)

    return predict_and_correct, iterate_correction


def run(state, n_steps, tol, max_iter, predict_and_correct, iterate_correction,
        correct_velocity, callback=None):
    """``__main__.py:125-203``: time loop with the corrector-count semantics.

    Returns ``(state, counts, errors)`` where ``counts[i]`` is the number of
    extra corrector iterations of step i (``c`` in ``__main__.py:164-186``)
    and ``errors[i]`` the final error of that step.  The loop stops after a
    step whose corrector did not converge (``__main__.py:128``); a NaN error
    also stops it (reported as such, the reference stops silently).
    """
    counts, errors = [], []
    last_error = 0.0
    it = 0
    while it < n_steps and last_error < tol:
        state_prev = state
        state, rhs, error = predict_and_correct(state_prev)
        c = 0
        while error > tol and c < max_iter:
            state, rhs, error = iterate_correction(state_prev, state, rhs)
            c += 1
        state = correct_velocity(state)
        counts.append(c)
        errors.append(error)
        last_error = error
        it += 1
        if callback is not None:
            callback(it, state)
        if math.isnan(error):
            break
    return state, counts, errors
