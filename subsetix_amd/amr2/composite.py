"""Sparse two-level composite: a full-domain coarse level and a fine level
that exists only on the refined region plus a one-cell ghost ring.

Step (the large single-instance stencil + transfer workload):
  1. prolong coarse -> fine ghost cells (parent values);
  2. upwind on the coarse level and on the fine level (ghosts included);
  3. restrict fine -> coarse on the refined cells (FP64 accumulation).
Every transfer is one fused pass (subsetix_amd.levels); fields are float32 or
float64, so a float32 run can be compared with a float64 run of the same
initial data (max / L2 error and mass drift).
"""

from __future__ import annotations

from dataclasses import dataclass

import torch

from ..expressions import IntervalSet, binary_op
from ..interval_field import IntervalField, create_interval_field
from ..interval_stencil import upwind_values
from ..levels import prolong_into, restrict_into
from ..morphology import clip_interval_set, full_interval_set, ghost_zones
from ..multilevel import prolong_set
from .mass import total_mass


@dataclass
class CompositeStats:
    coarse_cells: int
    fine_cells: int
    ghost_cells: int

    @property
    def cell_updates_per_step(self) -> int:
        return self.coarse_cells + self.fine_cells + self.ghost_cells


class CompositeTwoLevel:
    def __init__(self, width: int, height: int, refine: IntervalSet, *, ratio: int = 2, dtype=torch.float32,
                 velocity=(1.0, 0.5), cfl: float = 0.5, accumulate: str = "f64"):
        self.width, self.height, self.ratio = int(width), int(height), int(ratio)
        self.dtype = dtype
        self.velocity = (float(velocity[0]), float(velocity[1]))
        self.accumulate = accumulate
        self.coarse_set = full_interval_set(self.width, self.height, device=refine.device)
        self.refine = clip_interval_set(refine, width=self.width, height=self.height)
        self.fine = prolong_set(self.refine, self.ratio)
        fw, fh = self.width * self.ratio, self.height * self.ratio
        self.ghost = ghost_zones(self.fine, 1, 1, width=fw, height=fh)
        self.ext = binary_op(self.fine, self.ghost, "union")
        c0 = create_interval_field(self.coarse_set, 0.0, dtype=dtype)
        f0 = create_interval_field(self.ext, 0.0, dtype=dtype)
        self._c_co, self._f_co = c0.interval_cell_offsets, f0.interval_cell_offsets
        self._cbuf = [c0.values, torch.empty_like(c0.values)]
        self._fbuf = [f0.values, torch.empty_like(f0.values)]
        self._cur = 0
        self.dx_c = 1.0 / self.width
        self.dy_c = 1.0 / self.height
        self.dx_f = self.dx_c / self.ratio
        self.dy_f = self.dy_c / self.ratio
        a, b = self.velocity
        denom = abs(a) / self.dx_f + abs(b) / self.dy_f
        self.dt = cfl / denom if denom > 0 else 0.0

    @property
    def coarse(self) -> IntervalField:
        return IntervalField(self.coarse_set, self._cbuf[self._cur], self._c_co)

    @property
    def fine_field(self) -> IntervalField:
        return IntervalField(self.ext, self._fbuf[self._cur], self._f_co)

    def stats(self) -> CompositeStats:
        return CompositeStats(int(self.coarse.values.numel()), int(self.fine_field.values.numel())
                              - self._ghost_cells(), self._ghost_cells())

    def _ghost_cells(self) -> int:
        return int((self.ghost.end - self.ghost.begin).sum().item()) if self.ghost.interval_count else 0

    def set_values(self, coarse_values: torch.Tensor, fine_values: torch.Tensor | None = None) -> None:
        """Coarse values (W*H, row-major); fine values on the extended fine
        set (default: prolonged from the coarse level)."""
        self.coarse.values.copy_(coarse_values.to(self.dtype))
        if fine_values is None:
            prolong_into(self.ext, self.fine_field, self.coarse, self.ratio)
        else:
            self.fine_field.values.copy_(fine_values.to(self.dtype))
        restrict_into(self.refine, self.coarse, self.fine_field, self.ratio, accumulate=self.accumulate)
        prolong_into(self.ghost, self.fine_field, self.coarse, self.ratio)

    def prepare(self, tune: bool = False) -> None:
        """Build every cached table a step uses (cell offsets of the transfer
        targets, stencil segment tables), so that steps launch kernels only
        and can be captured into a graph.  ``tune``: also run the step's
        kernels once under the schedule autotuner (the state is left as it
        was: the ghost prolongation is idempotent and the other kernels write
        only the next buffers, which a step overwrites whole)."""
        from ..interval_field import cell_offsets
        from ..interval_stencil import stencil_plan

        from ..levels import transfer_plan

        for t in (self.ghost, self.refine):
            if t.interval_count:
                cell_offsets(t)
        if self.ratio == 2:
            if self.ghost.interval_count:
                transfer_plan(0, self.ghost, self.fine_field, self.coarse)
            if self.refine.interval_count:
                transfer_plan(1, self.refine, self.coarse, self.fine_field)
        stencil_plan(self.coarse_set, self._c_co, self.width)
        stencil_plan(self.ext, self._f_co, self.width * self.ratio)
        if tune:
            from .. import autotune

            with autotune.tuning():
                self._kernels(self._cur)

    def _kernels(self, cur: int) -> None:
        """The kernels of one step from buffer ``cur`` into ``1 - cur``."""
        a, b = self.velocity
        c = IntervalField(self.coarse_set, self._cbuf[cur], self._c_co)
        f = IntervalField(self.ext, self._fbuf[cur], self._f_co)
        nc = IntervalField(self.coarse_set, self._cbuf[1 - cur], self._c_co)
        nf = IntervalField(self.ext, self._fbuf[1 - cur], self._f_co)
        prolong_into(self.ghost, f, c, self.ratio)
        upwind_values(c, c.values, nc.values, width=self.width, height=self.height, a=a, b=b, dt=self.dt,
                      dx=self.dx_c, dy=self.dy_c, check_rows=False)
        upwind_values(f, f.values, nf.values, width=self.width * self.ratio, height=self.height * self.ratio, a=a,
                      b=b, dt=self.dt, dx=self.dx_f, dy=self.dy_f, check_rows=False)
        restrict_into(self.refine, nc, nf, self.ratio, accumulate=self.accumulate)

    def capture(self, stream: "torch.cuda.Stream", tune: bool = False) -> "torch.cuda.CUDAGraph":
        """Graph of two steps from the current buffer (ends on it again)."""
        from .._lib import context

        self.prepare(tune)
        stream.wait_stream(torch.cuda.current_stream(stream.device))
        with torch.cuda.stream(stream):
            context(stream.device)  # per-stream context allocated outside the capture
        g = torch.cuda.CUDAGraph()
        start = self._cur
        with torch.cuda.graph(g, stream=stream):
            self.step()
            self.step()
        self._cur = start
        torch.cuda.current_stream(stream.device).wait_stream(stream)
        return g

    def step(self) -> None:
        self._kernels(self._cur)
        self._cur = 1 - self._cur

    def mass(self) -> float:
        """FP64 composite mass (coarse level after the restriction)."""
        return total_mass(self.coarse.values, self.dx_c * self.dy_c)
