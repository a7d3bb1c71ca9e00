"""Plane Couette flow (wall-bounded, influence-matrix method) on the GPU.

Mirrors the flow-module surface of the reference's
``flows/plane_couette.py``: the callables ``_get_rhs``, ``_predict``,
``_correct``, ``_norm`` (``flows/plane_couette.py:274-549``), the
``make_stepper`` pair (``flows/plane_couette.py:552-572``), the pass-through
``correct_velocity`` (``flows/plane_couette.py:575-582``) and
``init_state``.  The influence-matrix tables (banded LU factors, pivots,
homogeneous solutions, M^-1) are built by the native context at creation.
"""

from __future__ import annotations

import torch

from ..fields import wall_field
from ..params import Params
from ..snapshots import load_snapshot
from .base import NativeFlow


class PlaneCouetteFlow(NativeFlow):
    def __init__(self, params: Params):
        if not params.walls:
            raise ValueError("PlaneCouetteFlow needs a walled system")
        super().__init__(params)

    def init_state(self, seed: int | None = None, rms: float = 0.0,
                   snapshot: str | None = None) -> torch.Tensor:
        """Laminar start (zeros), a snapshot's ``velocity_phys_nonexpanded``
        (``flows/plane_couette.py:225-248``), or the seeded no-slip field of
        ``dnsjax_amd.fields.wall_field``."""
        p = self.params
        if snapshot is not None:
            return load_snapshot(snapshot, p, self._native.device)
        if seed is None:
            return torch.zeros(self.spec_shape, dtype=torch.complex128, device=self._native.device)
        return self.to_device(wall_field(p.nx, p.ny, p.nz, p.lx, p.lz, rms, seed))
