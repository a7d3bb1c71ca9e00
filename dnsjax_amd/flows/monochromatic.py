"""Triply periodic flows (Kolmogorov, Waleffe, decaying box) on the GPU.

Mirrors the flow-module surface of the reference's
``flows/monochromatic.py``: the four stepper callables
(``_get_rhs``/``_predict``/``_correct``/``_norm``,
``flows/monochromatic.py:366-410``), ``predict_and_correct`` /
``iterate_correction`` built by ``make_stepper``
(``flows/monochromatic.py:413-427``), ``correct_velocity``
(``flows/monochromatic.py:545-572``) and ``init_state``
(``flows/monochromatic.py:294-314``).  The set-up tables (wavenumbers,
base flow, Helmholtz coefficients) live in the native context.
"""

from __future__ import annotations

import torch

from ..fields import periodic_field
from ..params import Params
from ..snapshots import load_snapshot
from .base import NativeFlow


class TriplyPeriodicFlow(NativeFlow):
    def __init__(self, params: Params):
        if params.walls:
            raise ValueError("TriplyPeriodicFlow needs a periodic system")
        super().__init__(params)

    def init_state(self, seed: int | None = None, rms: float = 0.0,
                   snapshot: str | None = None) -> torch.Tensor:
        """Laminar start (zeros, ``flows/monochromatic.py:296-301``), a
        snapshot's ``velocity_phys`` minus the base flow
        (``flows/monochromatic.py:302-311``), or the seeded solenoidal field
        of ``dnsjax_amd.fields.periodic_field``."""
        p = self.params
        if snapshot is not None:
            return load_snapshot(snapshot, p, self._native.device)
        if seed is None:
            return torch.zeros(self.spec_shape, dtype=torch.complex128, device=self._native.device)
        u = periodic_field(p.nx, p.ny, p.nz, p.lx, p.ly, p.lz, rms, seed)
        return self.to_device(u)
