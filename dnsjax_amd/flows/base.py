"""Shared flow object: the four stepper callables over one native context."""

from __future__ import annotations

import numpy as np
import torch

from ..native import NativeContext
from ..params import Params
from ..timestep import make_stepper


class NativeFlow:
    """Holds the device set-up of one flow and exposes the reference's
    callables and flow-module surface.

    ``get_rhs``, ``predict``, ``correct`` and ``norm`` have the signatures of
    the reference's ``get_rhs_fn``/``predict_fn``/``correct_fn``/``norm_fn``
    (``timestep.py:40-53``); ``predict_and_correct`` and
    ``iterate_correction`` come from ``make_stepper`` over them.
    """

    def __init__(self, params: Params):
        self.params = params
        self._native = NativeContext(params)
        self.predict_and_correct, self.iterate_correction = make_stepper(
            self.get_rhs, self.predict, self.correct, self.norm
        )

    @property
    def spec_shape(self):
        return self._native.shape

    # ------------------------------------------------------------ callables
    def get_rhs(self, state, *args):
        return self._native.get_rhs(state)

    def predict(self, state, rhs_no_lapl, *args):
        return self._native.predict(state, rhs_no_lapl)

    def correct(self, state_prev, prediction_state, rhs_prev, rhs_next, *args):
        return self._native.correct(state_prev, prediction_state, rhs_prev, rhs_next)

    def norm(self, correction, *args):
        return self._native.norm(correction)

    # ------------------------------------------------------------ surface
    def to_device(self, state) -> torch.Tensor:
        if isinstance(state, np.ndarray):
            state = torch.from_numpy(np.ascontiguousarray(state, dtype=np.complex128))
        return state.to(self._native.device).contiguous()

    def save_snapshot(self, path: str, state) -> None:
        """Write ``state`` in the reference's snapshot format (``.npz``:
        periodic ``velocity_phys``, the total velocity on the padded grid;
        walls ``velocity_phys_nonexpanded``)."""
        from ..snapshots import save_snapshot

        save_snapshot(path, state, self.params)

    def get_stats(self, state) -> dict:
        """E, I, D, E' as 0-d device tensors (``flows/monochromatic.py:508-510``,
        ``flows/plane_couette.py:616-618``)."""
        v = self._native.get_stats(state)
        keys = ("E", "I", "D", "E'") if not self.params.walls else ("E'",)
        idx = {"E": 0, "I": 1, "D": 2, "E'": 3}
        return {k: v[idx[k]] for k in keys}

    def escalate(self, on: bool) -> None:
        """Mixed form: the following corrector passes run in FP64 (or back)."""
        self._native.escalate(on)

    def correct_velocity(self, state):
        """In place (the reference donates its input); returns
        ``(state, norm_corrections)``."""
        measure = self.params.measure_corrections
        state, err = self._native.correct_velocity(state, measure=measure and not self.params.walls)
        if not measure:
            return state, None
        return state, ({"div": err} if err is not None else {})
