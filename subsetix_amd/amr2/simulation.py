"""Imperative two-level AMR advection driver (drop-in for
``subsetix_amr2.simulation``, simulation.py:106-354).

The step order is the reference's: synchronise -> regrid every k steps ->
advance both levels -> synchronise.  Both levels are full-domain interval
fields (coarse W x H, fine rW x rH) as in the reference; the refine set only
selects where the synchronisation writes.

Extensions (keyword-only, defaults reproduce the reference):
  * ``SimulationConfig.tag`` = "gradient" (reference) or "value" (tag cells
    whose value is >= refine_threshold);
  * float64 fields (the reference is float32-only);
  * ``SimulationConfig.accumulate`` = "f64" sums restrictions in FP64;
  * ``total_mass()``: FP64 mass of the coarse level (after synchronisation
    it carries the composite mass).
"""

from __future__ import annotations

import math
from dataclasses import dataclass
from typing import Callable, Iterable, Sequence

import torch

from ..expressions import IntervalSet
from ..interval_field import IntervalField, create_interval_field
from ..interval_stencil import step_upwind_interval_field
from ..morphology import full_interval_set
from ..multilevel import prolong_field
from .fields import Action, ActionField, synchronize_interval_fields
from .geometry import TwoLevelGeometry
from .mass import total_mass
from .regrid import enforce_two_level_grading_set, gradient_tag_threshold_interval_field, \
    intervals_above_threshold_interval_field


@dataclass(frozen=True)
class SquareSpec:
    center: tuple[float, float]
    half_width: tuple[float, float]
    value: float = 1.0


def build_square_interval_field(width: int, height: int, squares: Sequence[SquareSpec], *,
                                dtype=torch.float32) -> IntervalField:
    """simulation.py:57-103."""
    if width <= 0 or height <= 0:
        raise ValueError("width and height must be positive")
    specs = list(squares)
    if not specs:
        raise ValueError("at least one SquareSpec is required")
    interval = full_interval_set(width, height)
    field = create_interval_field(interval, fill_value=0.0, dtype=dtype)
    grid = field.values.view(height, width)
    x_scale = width - 1 if width > 1 else 1
    y_scale = height - 1 if height > 1 else 1
    for spec in specs:
        cx, cy = spec.center
        hx, hy = spec.half_width
        x0 = max(0, min(width - 1, math.ceil((cx - hx) * x_scale)))
        x1 = max(0, min(width - 1, math.floor((cx + hx) * x_scale)))
        y0 = max(0, min(height - 1, math.ceil((cy - hy) * y_scale)))
        y1 = max(0, min(height - 1, math.floor((cy + hy) * y_scale)))
        if x0 > x1 or y0 > y1:
            continue
        value = float(spec.value)
        if value == 0.0:
            continue
        grid[y0:y1 + 1, x0:x1 + 1] = value
    return field


_DEFAULT_SQUARES = [
    SquareSpec(center=(0.3, 0.3), half_width=(0.1, 0.1), value=1.0),
    SquareSpec(center=(0.7, 0.7), half_width=(0.1, 0.1), value=1.0),
]


@dataclass(frozen=True)
class SimulationConfig:
    coarse_resolution: int
    velocity: tuple[float, float]
    cfl: float = 0.9
    refine_threshold: float = 0.05
    grading: int = 1
    grading_mode: str = "von_neumann"
    ratio: int = 2
    tag: str = "gradient"
    accumulate: str = "f32"


@dataclass
class AMRState:
    coarse_field: IntervalField
    fine_field: IntervalField
    actions: ActionField
    geometry: TwoLevelGeometry

    @property
    def coarse(self) -> torch.Tensor:
        return self.coarse_field.values.view(self.geometry.height, self.geometry.width)

    @property
    def fine(self) -> torch.Tensor:
        r = self.geometry.ratio
        return self.fine_field.values.view(self.geometry.height * r, self.geometry.width * r)


@dataclass(frozen=True)
class SimulationStats:
    step: int
    time: float
    dt: float
    dx_coarse: float
    dy_coarse: float
    ratio: int


Callback = Callable[[AMRState, SimulationStats], None]


class AMR2Simulation:
    def __init__(self, config: SimulationConfig):
        self.config = config
        self._validate_config()
        self.width = int(config.coarse_resolution)
        self.height = self.width
        self.dx_coarse = 1.0 / self.width
        self.dy_coarse = 1.0 / self.height
        self.dx_fine = self.dx_coarse / config.ratio
        self.dy_fine = self.dy_coarse / config.ratio
        denom = abs(config.velocity[0]) / self.dx_fine + abs(config.velocity[1]) / self.dy_fine
        self.dt = config.cfl / denom if denom > 0 else 0.0
        self.state: AMRState | None = None
        self.current_step = 0
        self.current_time = 0.0

    def initialize_square(self, *, squares: Sequence[SquareSpec] | None = None, amplitude: float = 1.0,
                          dtype=torch.float32) -> None:
        specs = list(squares) if squares is not None else _DEFAULT_SQUARES
        field = build_square_interval_field(self.width, self.height, specs, dtype=dtype)
        if amplitude != 1.0:
            field.values.mul_(float(amplitude))
        self.set_initial_field(field)

    def set_initial_field(self, coarse_field: IntervalField) -> None:
        """simulation.py:201-230 (validations included)."""
        if not isinstance(coarse_field, IntervalField):
            raise TypeError("coarse_field must be an IntervalField")
        h, w = self.height, self.width
        s = coarse_field.interval_set
        if s.row_count != h:
            raise ValueError("coarse_field height mismatch with simulation domain")
        rows_ok = s.is_dense_rows()
        if rows_ok is None:
            from .. import _core

            rows_ok = bool(torch.equal(s.rows, _core.arange_i32(h, s.device)))
        if not rows_ok:
            raise ValueError("coarse_field rows must match the simulation row ids [0, height)")
        if s.interval_count != h:
            raise ValueError("coarse_field must contain exactly one interval per row")
        if not (bool((s.begin == 0).all()) and bool((s.end == w).all())):
            raise ValueError("coarse_field must cover the full width on every row")
        if int(coarse_field.values.numel()) != w * h:
            raise ValueError("coarse_field must hold width * height cells")
        refine = self._refine_from_field(coarse_field)
        actions = ActionField.full_grid(h, w, self.config.ratio, default=Action.KEEP)
        actions.set_from_interval_set(refine)
        geometry = TwoLevelGeometry.from_action_field(actions)
        fine_field = prolong_field(coarse_field, self.config.ratio, dtype=coarse_field.values.dtype)
        self.state = AMRState(coarse_field=coarse_field, fine_field=fine_field, actions=actions, geometry=geometry)
        self.current_step = 0
        self.current_time = 0.0

    def run(self, *, steps: int, regrid_every: int = 1, exporters: Iterable[Callback] | None = None,
            listeners: Iterable[Callback] | None = None):
        state = self._require_state()
        exporters_l = list(exporters or [])
        listeners_l = list(listeners or [])
        stats = self._build_stats(self.current_step)
        for cb in exporters_l + listeners_l:
            cb(state, stats)
        period = max(1, int(regrid_every))
        for _ in range(int(steps)):
            self._synchronize()
            if self.current_step % period == 0:
                self._update_geometry()
            self._advance()
            self._synchronize()
            self.current_step += 1
            self.current_time += self.dt
            state = self._require_state()
            stats = self._build_stats(self.current_step)
            for cb in exporters_l:
                cb(state, stats)
            for cb in listeners_l:
                cb(state, stats)
        return state, stats

    def total_mass(self) -> float:
        state = self._require_state()
        return total_mass(state.coarse_field.values, self.dx_coarse * self.dy_coarse)

    # -- internals -------------------------------------------------------
    def _validate_config(self) -> None:
        c = self.config
        if c.coarse_resolution <= 0:
            raise ValueError("coarse_resolution must be positive")
        if c.ratio < 1:
            raise ValueError("ratio must be >= 1")
        if c.grading_mode not in {"von_neumann", "moore"}:
            raise ValueError("grading_mode must be 'von_neumann' or 'moore'")
        if c.refine_threshold <= 0.0:
            raise ValueError("refine_threshold must be > 0")
        if c.tag not in {"gradient", "value"}:
            raise ValueError("tag must be 'gradient' or 'value'")

    def _require_state(self) -> AMRState:
        if self.state is None:
            raise RuntimeError("simulation not initialised")
        return self.state

    def _refine_from_field(self, field: IntervalField) -> IntervalSet:
        if self.config.tag == "gradient":
            tagged = gradient_tag_threshold_interval_field(field, width=self.width, height=self.height,
                                                           threshold=self.config.refine_threshold)
        else:
            tagged = intervals_above_threshold_interval_field(field, width=self.width, height=self.height,
                                                              threshold=self.config.refine_threshold)
        return enforce_two_level_grading_set(tagged, padding=self.config.grading, mode=self.config.grading_mode,
                                             width=self.width, height=self.height)

    _refine_from_gradient = _refine_from_field

    def _synchronize(self) -> None:
        s = self._require_state()
        synchronize_interval_fields(s.coarse_field, s.fine_field, s.actions, ratio=self.config.ratio, reducer="mean",
                                    fill_fine_outside=True, copy=False, accumulate=self.config.accumulate)

    def _update_geometry(self) -> None:
        s = self._require_state()
        refine = self._refine_from_field(s.coarse_field)
        s.actions.set_from_interval_set(refine)
        s.geometry = TwoLevelGeometry.from_action_field(s.actions, workspace=s.geometry.workspace)

    def _advance(self) -> None:
        s = self._require_state()
        a, b = self.config.velocity
        g = s.geometry
        r = g.ratio
        cdt = s.coarse_field.values.dtype
        s.coarse_field = step_upwind_interval_field(s.coarse_field, width=g.width, height=g.height, a=a, b=b, dt=self.dt,
                                                    dx=self.dx_coarse, dy=self.dy_coarse, compute_dtype=cdt)
        s.fine_field = step_upwind_interval_field(s.fine_field, width=g.width * r, height=g.height * r, a=a, b=b,
                                                  dt=self.dt, dx=self.dx_fine, dy=self.dy_fine, compute_dtype=cdt)

    def _build_stats(self, step: int) -> SimulationStats:
        return SimulationStats(step=step, time=self.current_time, dt=self.dt, dx_coarse=self.dx_coarse,
                               dy_coarse=self.dy_coarse, ratio=self.config.ratio)


__all__ = [
    "AMR2Simulation",
    "AMRState",
    "SimulationConfig",
    "SimulationStats",
    "SquareSpec",
    "build_square_interval_field",
]
