"""Run parameters: the subset of the reference's ``parameters.py`` the
time-step path reads, as one flat dataclass.

Field meanings follow ``parameters.py:26-133`` (sections phys, geo, res,
step, debug).  ``ly`` is a parameter here; the reference fixes it to 4 for
periodic systems and 2 for walls (``parameters.py:186-193``), which is
what ``None`` selects.  ``precision`` selects the plain FP64 path or the
mixed form (FP32 transforms and nonlinear product).
"""

from __future__ import annotations

import math
from dataclasses import dataclass, replace

PERIODIC_SYSTEMS = ("decaying-box", "kolmogorov", "waleffe")  # parameters.py:19-20
WALLED_SYSTEMS = ("plane-couette",)  # parameters.py:22-23


@dataclass
class Params:
    system: str = "kolmogorov"
    nx: int = 128
    ny: int = 128
    nz: int = 128
    lx: float = 4.0
    ly: float | None = None
    lz: float = 4.0
    re: float = 1000.0
    oversampling_factor: int = 3
    oversample_y: bool = True
    tilt_degree: float = 0.0
    fd_order: int = 4
    dt: float = 0.01
    implicitness: float = 0.5
    corrector_tolerance: float = 1e-5
    max_corrector_iterations: int = 10
    correct_divergence: bool = True
    measure_corrections: bool = False
    precision: str = "fp64"
    tile_cols: int = 0  # 0: autotune kernel variants; else use `variants`
    variants: int = 0
    # Mixed form only.  Every corrector pass of a step is evaluated as an
    # FP32 increment over the step's first evaluation (split precision), so
    # its rounding scales with the increment and the corrector converges to
    # corrector_tolerance as in FP64.  plain_mixed evaluates each pass in
    # full in FP32 instead (its corrector error stalls near 1e-10 at 128^3).
    plain_mixed: bool = False
    # Mixed form only: escalation to FP64 passes.  A pass whose error, at or
    # below `escalate_cap_value`, is not below escalate_ratio times the error
    # two passes earlier has reached the resolution of its FP32 evaluations;
    # the step's remaining passes then run in FP64, so every step still ends
    # at error <= corrector_tolerance (__main__.py:167-169).  None disables it.
    escalate_ratio: float | None = None
    escalate_cap: float = 1e4  # times corrector_tolerance
    # Absolute ceiling of that window: a stall above 1e-8 is slow contraction
    # (the wall-bounded corrector repeats its error on the pass after a
    # start), not FP32 resolution.
    escalate_abs: float = 1e-8
    # Split evaluations only: re-centring of the reference.  The rounding of
    # an increment scales with its size, and the increments are measured
    # from the step's first corrector iterate; a step whose predict-and-
    # correct error exceeds recenter_far times corrector_tolerance records a
    # new reference at its first pass that starts from an error of at most
    # recenter_near times the tolerance, so its later increments are small
    # again (early: the new reference's own FP32 rounding then shows in a
    # pass whose error is far above it).  None disables it; the default is on for wall-bounded flows,
    # whose first corrections grow with the perturbation (at 128 x 256 x 128
    # to 1e-3 and beyond, which put the split floor at the 1e-10 tolerance),
    # and off for periodic ones (whose floor stays orders below it).
    recenter_far: float | None = -1.0
    recenter_near: float = 5e5
    # Wall-bounded flows: build the influence-matrix tables (band LU,
    # homogeneous solutions, M^-1, chunk maps) on the host instead of the
    # GPU (the two agree bit for bit; the GPU set-up falls back to the host
    # where a band does not fit its LDS).
    setup_on_host: bool = False

    def __post_init__(self) -> None:
        if self.system not in PERIODIC_SYSTEMS + WALLED_SYSTEMS:
            raise ValueError(f"unknown system {self.system!r}")
        if self.ly is None:
            self.ly = 4.0 if self.system in PERIODIC_SYSTEMS else 2.0
        if self.system in WALLED_SYSTEMS and self.ly != 2.0:
            # parameters.py:190-191: the walls sit at y = -1 and y = 1
            raise ValueError("walled systems have ly = 2 (y in [-1, 1])")
        if self.precision not in ("fp64", "mixed"):
            raise ValueError("precision must be 'fp64' or 'mixed'")
        if self.escalate_ratio is None and self.precision == "mixed":
            self.escalate_ratio = 0.5
        if self.recenter_far is not None and self.recenter_far < 0:  # the default
            self.recenter_far = 1e6 if self.system in WALLED_SYSTEMS else None
        if self.recenter_far is not None and not self.recenter_near < self.recenter_far:
            raise ValueError("need recenter_near < recenter_far")
        if not self.re > 0 or not self.dt > 0 or not 0 <= self.implicitness <= 1:
            raise ValueError("need re > 0, dt > 0, 0 <= implicitness <= 1")

    @property
    def escalate_cap_value(self) -> float:
        """Error at or below which a stalled pass escalates to FP64."""
        return min(self.escalate_cap * self.corrector_tolerance, self.escalate_abs)

    @property
    def recenter_far_value(self) -> float:
        """Predict-and-correct error above which a step re-centres its split reference."""
        return math.inf if self.recenter_far is None else self.recenter_far * self.corrector_tolerance

    @property
    def recenter_near_value(self) -> float:
        return self.recenter_near * self.corrector_tolerance

    @property
    def walls(self) -> bool:
        return self.system in WALLED_SYSTEMS

    @property
    def spec_shape(self) -> tuple:
        """``sharding.py:395-434``."""
        if self.walls:
            return (self.nz - 1, self.nx // 2, self.ny)
        return (self.ny - 1, self.nz - 1, self.nx // 2)

    @property
    def padded(self) -> tuple:
        """(Py, Pz, Px), ``parameters.py:233-254``."""
        f = self.oversampling_factor
        py = (f * self.ny // 2 if self.oversample_y else self.ny) if not self.walls else self.ny
        return (py, f * self.nz // 2, f * self.nx // 2)

    def with_(self, **kw) -> "Params":
        return replace(self, **kw)


def kolmogorov_2pi(n: int, re: float, dt: float, **kw) -> Params:
    """A cubic 2 pi box Kolmogorov run (the benchmark boxes)."""
    L = 2 * math.pi
    return Params(system="kolmogorov", nx=n, ny=n, nz=n, lx=L, ly=L, lz=L, re=re, dt=dt, **kw)
