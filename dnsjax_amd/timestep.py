"""Predictor-corrector stepper factory: the ``make_stepper`` contract.

Same signature and return values as the reference's ``timestep.make_stepper``
(``timestep.py:23-122``): four callables in, ``(predict_and_correct,
iterate_correction)`` out, with

    predict_and_correct(state, *args) -> (prediction, rhs_next, error)
    iterate_correction(state_prev, prediction, rhs_prev, *args)
        -> (prediction_next, rhs_next, error)

``predict_and_correct`` leaves ``state`` untouched; ``iterate_correction``
consumes ``prediction`` (the reference donates it, ``timestep.py:97``: here
it is updated in place and returned).  ``error`` is a 0-d float64 device
tensor, compared on the host by the caller as in ``__main__.py:167``.

When all four callables are the bound methods of one native flow (objects
exposing ``_native``), the returned functions are the fused device entry
points of libdnsmi (one RHS pipeline with the predictor, corrector and norm
in its last pass) -- the counterpart of the reference's ``jit`` of the
composition.  Otherwise the callables are composed in Python exactly as the
reference composes them.
"""

from __future__ import annotations

from collections.abc import Callable


def _common_native(fns):
    owners = {getattr(f, "__self__", None) for f in fns}
    if len(owners) != 1:
        return None
    owner = owners.pop()
    nat = getattr(owner, "_native", None)
    if nat is None:
        return None
    names = [getattr(f, "__name__", "") for f in fns]
    if names != ["get_rhs", "predict", "correct", "norm"]:
        return None
    return nat


def make_stepper(
    get_rhs_fn: Callable,
    predict_fn: Callable,
    correct_fn: Callable,
    norm_fn: Callable,
):
    nat = _common_native((get_rhs_fn, predict_fn, correct_fn, norm_fn))
    if nat is not None:

        def predict_and_correct(state, *args):
            return nat.predict_and_correct(state)

        def iterate_correction(state_prev, prediction_state, rhs_prev, *args):
            return nat.iterate_correction(state_prev, prediction_state, rhs_prev)

        predict_and_correct.fused = iterate_correction.fused = True
        return predict_and_correct, iterate_correction

    def predict_and_correct(state, *args):  # timestep.py:70-94
        rhs_prev = get_rhs_fn(state, *args)
        prediction_state = predict_fn(state, rhs_prev, *args)
        rhs_next = get_rhs_fn(prediction_state, *args)
        prediction_state, correction = correct_fn(state, prediction_state, rhs_prev, rhs_next, *args)
        error = norm_fn(correction, *args)
        return prediction_state, rhs_next, error

    def iterate_correction(state_prev, prediction_state, rhs_prev, *args):  # timestep.py:98-120
        rhs_next = get_rhs_fn(prediction_state, *args)
        prediction_state, correction = correct_fn(state_prev, prediction_state, rhs_prev, rhs_next, *args)
        error = norm_fn(correction, *args)
        return prediction_state, rhs_next, error

    predict_and_correct.fused = iterate_correction.fused = False
    return predict_and_correct, iterate_correction
