"""Imperative two-level AMR advection driver (drop-in for
``subsetix_amr2.simulation``; public names and step order of
simulation.py:106-354).

Both levels are full-domain fields (coarse W x H, fine rW x rH) as in the
reference, and a step is synchronise -> regrid every k steps -> advance both
levels -> synchronise.  The machinery under that surface is this package's:

* the two level fields live in a :class:`_LevelPair` of ping-pong buffers,
  so a step allocates nothing and advances each level with one stencil
  launch (segment tables built once per level, since the full-domain
  geometry never changes);
* synchronisation is two fused level transfers (``levels.restrict_into`` on
  the refine set, ``levels.prolong_into`` on the fine cells outside it)
  instead of the restrict-whole-field / gather / scatter chain of
  fields.py:570-612; cell for cell the result is the same (mean of the
  children in dy-major order, parent copy), and the transfer tables are
  rebuilt only when the refine set changes;
* the initial square field is assembled on the host as one array and copied
  to the device once.

Extensions (keyword-only; the defaults reproduce the reference):
  * ``SimulationConfig.tag``: "gradient" (reference) or "value" (cells whose
    value is >= refine_threshold);
  * float64 fields (the reference is float32-only);
  * ``SimulationConfig.accumulate = "f64"`` sums the restriction in FP64;
  * ``total_mass()``: FP64 mass of the coarse level, which carries the
    composite mass after synchronisation.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Callable, Dict, Iterable, List, Sequence

import numpy as np
import torch

from ..expressions import IntervalSet
from ..interval_field import IntervalField, create_interval_field
from ..interval_stencil import upwind_values
from ..levels import prolong_into
from ..morphology import full_interval_set
from ..multilevel import prolong_set
from .fields import Action, ActionField, sync_levels
from .geometry import TwoLevelGeometry
from .mass import total_mass
from .regrid import enforce_two_level_grading_set, gradient_tag_threshold_interval_field, \
    intervals_above_threshold_interval_field


@dataclass(frozen=True)
class SquareSpec:
    """A square patch [cx - hx, cx + hx] x [cy - hy, cy + hy] in unit
    coordinates, filled with ``value``."""

    center: tuple[float, float]
    half_width: tuple[float, float]
    value: float = 1.0


def _covered_indices(n: int, lo: float, hi: float) -> np.ndarray:
    """Mask of the indices i in [0, n) with lo <= i <= hi, where the unit
    coordinate of index i is i / (n - 1) (i for n = 1)."""
    scale = float(n - 1 if n > 1 else 1)
    idx = np.arange(n, dtype=np.float64)
    return (idx >= lo * scale) & (idx <= hi * scale)


def build_square_interval_field(width: int, height: int, squares: Sequence[SquareSpec], *,
                                dtype=torch.float32) -> IntervalField:
    """Full-domain coarse field that holds each square's value on the cells
    it covers (simulation.py:57-103 semantics: cell (i, j) is covered when its
    index lies in the square scaled by (width - 1, height - 1); later squares
    overwrite earlier ones; zero-valued squares are skipped)."""
    w, h = int(width), int(height)
    if w <= 0 or h <= 0:
        raise ValueError(f"the field needs a positive size, got width={width}, height={height}")
    specs = list(squares)
    if not specs:
        raise ValueError("no squares given: pass at least one SquareSpec")
    grid = np.zeros((h, w), dtype=np.float64)
    for sq in specs:
        if float(sq.value) == 0.0:
            continue
        mx = _covered_indices(w, sq.center[0] - sq.half_width[0], sq.center[0] + sq.half_width[0])
        my = _covered_indices(h, sq.center[1] - sq.half_width[1], sq.center[1] + sq.half_width[1])
        grid[np.ix_(my, mx)] = float(sq.value)
    field = create_interval_field(full_interval_set(w, h), fill_value=0.0, dtype=dtype)
    field.values.copy_(torch.from_numpy(grid.reshape(-1)).to(field.values.dtype))
    return field


# simulation.py:403-405: one square centred at (0.3, 0.3)
_DEFAULT_SQUARES = (SquareSpec(center=(0.30, 0.30), half_width=(0.10, 0.10)),)


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

_CONFIG_RULES = (
    (lambda c: c.coarse_resolution > 0, "coarse_resolution"),
    (lambda c: c.ratio >= 1, "ratio"),
    (lambda c: c.grading_mode in ("von_neumann", "moore"), "grading_mode"),
    (lambda c: c.refine_threshold > 0.0, "refine_threshold"),
    (lambda c: c.tag in ("gradient", "value"), "tag"),
    (lambda c: c.accumulate in ("f32", "f64"), "accumulate"),
)


class _LevelPair:
    """The coarse and fine full-domain fields of one run, each as two
    buffers (current / next) over one cached geometry."""

    def __init__(self, coarse: IntervalField, ratio: int, width: int, height: int):
        self.ratio, self.width, self.height = ratio, width, height
        dt = coarse.values.dtype
        fine = create_interval_field(full_interval_set(width * ratio, height * ratio, device=coarse.values.device),
                                     0.0, dtype=dt)
        self.sets = (coarse.interval_set, fine.interval_set)
        self.offsets = (coarse.interval_cell_offsets, fine.interval_cell_offsets)
        self.bufs = ([coarse.values, torch.empty_like(coarse.values)], [fine.values, torch.empty_like(fine.values)])
        self.cur = 0
        # the fine level starts as the parents' values on every cell
        prolong_into(fine.interval_set, self.field(1), self.field(0), ratio)

    def field(self, level: int, which: int | None = None) -> IntervalField:
        k = self.cur if which is None else which
        return IntervalField(self.sets[level], self.bufs[level][k], self.offsets[level])

    def advance(self, a: float, b: float, dt: float, spacing: List[tuple]) -> None:
        nxt = 1 - self.cur
        for lv in (0, 1):
            w = self.width * self.ratio ** lv
            h = self.height * self.ratio ** lv
            dx, dy = spacing[lv]
            upwind_values(self.field(lv), self.bufs[lv][self.cur], self.bufs[lv][nxt], width=w, height=h, a=a, b=b,
                          dt=dt, dx=dx, dy=dy, check_rows=False)
        self.cur = nxt


class AMR2Simulation:
    def __init__(self, config: SimulationConfig):
        broken = [name for ok, name in _CONFIG_RULES if not ok(config)]
        if broken:
            raise ValueError(f"invalid SimulationConfig field(s) {', '.join(broken)}: resolution > 0, ratio >= 1, "
                             "grading_mode von_neumann|moore, refine_threshold > 0, tag gradient|value, "
                             "accumulate f32|f64")
        self.config = config
        self.width = self.height = int(config.coarse_resolution)
        self.dx_coarse = 1.0 / self.width
        self.dy_coarse = 1.0 / self.height
        self.dx_fine = self.dx_coarse / config.ratio
        self.dy_fine = self.dy_coarse / config.ratio
        rate = abs(config.velocity[0]) / self.dx_fine + abs(config.velocity[1]) / self.dy_fine
        self.dt = config.cfl / rate if rate > 0 else 0.0
        self.state: AMRState | None = None
        self.current_step = 0
        self.current_time = 0.0
        self._levels: _LevelPair | None = None
        self._outside: IntervalSet | None = None  # fine cells not under the refine set

    # ------------------------------------------------------------ set-up
    def initialize_square(self, *, squares: Sequence[SquareSpec] | None = None, amplitude: float = 1.0,
                          dtype=torch.float32) -> None:
        field = build_square_interval_field(self.width, self.height,
                                            list(squares) if squares is not None else _DEFAULT_SQUARES, dtype=dtype)
        if amplitude != 1.0:
            field.values.mul_(float(amplitude))  # in the field's dtype, as simulation.py:183-185
        self.set_initial_field(field)

    def set_initial_field(self, coarse_field: IntervalField) -> None:
        """Attach a full-domain coarse field (one interval [0, W) on each row
        0..H-1), prolong it onto the fine level and regrid (simulation.py:
        201-230)."""
        if not isinstance(coarse_field, IntervalField):
            raise TypeError(f"set_initial_field expects an IntervalField, got {type(coarse_field).__name__}")
        self._require_full_domain(coarse_field)
        levels = _LevelPair(coarse_field, self.config.ratio, self.width, self.height)
        actions = ActionField.full_grid(self.height, self.width, self.config.ratio, default=Action.KEEP)
        self._levels = levels
        self.state = AMRState(coarse_field=levels.field(0), fine_field=levels.field(1), actions=actions,
                              geometry=None)  # type: ignore[arg-type]
        self._regrid()
        self.current_step = 0
        self.current_time = 0.0

    def _require_full_domain(self, f: IntervalField) -> None:
        s = f.interval_set
        full = full_interval_set(self.width, self.height, device=f.values.device)
        same = (s.row_count == full.row_count and s.interval_count == full.interval_count
                and int(f.values.numel()) == self.width * self.height)
        if same:
            same = all(torch.equal(x, y) for x, y in ((s.rows, full.rows), (s.row_offsets, full.row_offsets),
                                                      (s.begin, full.begin), (s.end, full.end)))
        if not same:
            raise ValueError(f"the coarse field must cover the whole {self.width} x {self.height} domain: rows "
                             f"0..{self.height - 1}, each holding the single interval [0, {self.width})")

    # -------------------------------------------------------------- loop
    def run(self, *, steps: int, regrid_every: int = 1, exporters: Iterable[Callback] | None = None,
            listeners: Iterable[Callback] | None = None):
        callbacks = list(exporters or []) + list(listeners or [])
        self._notify(callbacks)
        period = max(1, int(regrid_every))
        for _ in range(int(steps)):
            self._synchronize()
            if self.current_step % period == 0:
                self._regrid()
            a, b = self.config.velocity
            self._levels.advance(a, b, self.dt, [(self.dx_coarse, self.dy_coarse), (self.dx_fine, self.dy_fine)])
            self._publish()
            self._synchronize()
            self.current_step += 1
            self.current_time += self.dt
            self._notify(callbacks)
        return self.state, self._stats()

    def total_mass(self) -> float:
        st = self._state()
        return total_mass(st.coarse_field.values, self.dx_coarse * self.dy_coarse)

    # --------------------------------------------------------- internals
    def _state(self) -> AMRState:
        if self.state is None:
            raise RuntimeError("no state yet: call initialize_square() or set_initial_field() first")
        return self.state

    def _stats(self) -> SimulationStats:
        return SimulationStats(step=self.current_step, time=self.current_time, dt=self.dt,
                               dx_coarse=self.dx_coarse, dy_coarse=self.dy_coarse, ratio=self.config.ratio)

    def _notify(self, callbacks: Sequence[Callback]) -> None:
        if callbacks:
            st, stats = self._state(), self._stats()
            for cb in callbacks:
                cb(st, stats)

    def _publish(self) -> None:
        st = self._state()
        st.coarse_field = self._levels.field(0)
        st.fine_field = self._levels.field(1)

    def _tagged(self, coarse: IntervalField) -> IntervalSet:
        c = self.config
        tag = gradient_tag_threshold_interval_field if c.tag == "gradient" else \
            intervals_above_threshold_interval_field
        tagged = tag(coarse, width=self.width, height=self.height, threshold=c.refine_threshold)
        return enforce_two_level_grading_set(tagged, padding=c.grading, mode=c.grading_mode, width=self.width,
                                             height=self.height)

    def _regrid(self) -> None:
        """Tag + grade the coarse field, then refresh the action field, the
        geometry and the fine-level target outside the refine set."""
        st = self._state()
        refine = self._tagged(st.coarse_field)
        st.actions.set_from_interval_set(refine)
        ws = st.geometry.workspace if st.geometry is not None else None
        st.geometry = TwoLevelGeometry.from_action_field(st.actions, workspace=ws)
        self._outside = prolong_set(st.geometry.coarse_only, self.config.ratio)

    def _synchronize(self) -> None:
        """Coarse cells under the refine set take the mean of their children;
        fine cells outside it take their parent's value."""
        st = self._state()
        sync_levels(self._levels.field(0), self._levels.field(1), st.actions.refine_set(), self._outside,
                    ratio=self.config.ratio, accumulate=self.config.accumulate)


class TwoLevelVTKExporter:
    """Exporter callback writing the state to VTK every ``every`` steps
    (simulation.py:360-400 interface; the writer is amr2.export)."""

    def __init__(self, directory: str, *, prefix: str = "amr2", every: int = 1, ghost_halo: int = 0):
        self.directory = directory
        self.prefix = prefix
        self.every = max(1, int(every))
        self.ghost_halo = max(0, int(ghost_halo))

    def __call__(self, state: AMRState, stats: SimulationStats) -> Dict[str, str] | None:
        if stats.step % self.every:
            return None
        from .export import save_two_level_vtk

        g = state.geometry
        return save_two_level_vtk(self.directory, self.prefix, stats.step, coarse_field=state.coarse_field,
                                  fine_field=state.fine_field, refine_set=state.actions.refine_set(),
                                  coarse_only_set=g.coarse_only, fine_set=g.fine, dx_coarse=stats.dx_coarse,
                                  dy_coarse=stats.dy_coarse, ratio=stats.ratio, time_value=stats.time,
                                  ghost_halo=self.ghost_halo, width=g.width, height=g.height)


__all__ = [
    "AMR2Simulation",
    "AMRState",
    "SimulationConfig",
    "SimulationStats",
    "SquareSpec",
    "TwoLevelVTKExporter",
    "build_square_interval_field",
]
