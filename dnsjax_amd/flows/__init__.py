"""Flow modules: the flow-module surface the reference's driver consumes
(``__main__.py:56-76``): ``init_state``, ``predict_and_correct``,
``iterate_correction``, ``correct_velocity``, ``get_stats``."""

from .base import NativeFlow
from .monochromatic import TriplyPeriodicFlow
from .plane_couette import PlaneCouetteFlow


def make_flow(params):
    return PlaneCouetteFlow(params) if params.walls else TriplyPeriodicFlow(params)


__all__ = ["NativeFlow", "TriplyPeriodicFlow", "PlaneCouetteFlow", "make_flow"]
