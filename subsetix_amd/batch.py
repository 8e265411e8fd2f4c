"""Batched multi-level AMR: many independent advection instances advanced by
one launch per kernel and per step.

Layout ("stack"): instance i owns the row ids [base_L[i], base_L[i] + N_i R^L)
at level L, with base_{L+1} = R * base_L and a gap of empty rows between
instances, so one CSR set per level holds every instance.  Every row-local
kernel (set algebra, dilation, tagging, transfers, stencil) then serves the
whole batch in one launch: rows of different instances never interact,
floor(row / R) maps an instance's rows onto its own coarser rows, and the
per-row instance index gives the stencil its width and coefficients.

Levels: level 0 is the full N_i x N_i domain of every instance; levels 1 and
2 exist only on the refined regions plus a one-cell ghost ring.  A step is
  prolong ghosts (L0 -> L1, L1 -> L2), upwind on every level, restrict
  (L2 -> L1, L1 -> L0, FP64 sums)
and is captured once per regrid into a hipGraph (two steps, ping-pong
buffers) that is replayed until the next regrid.  Regrid (every k steps):
gradient tag on L0, grading dilation, one-cell L1 ring forced onto its
parents, gradient tag on L1 gated by the eroded L1 (3-level nesting:
L2 inside L1, no 8-neighbour contact between L2 and cells outside L1,
coarse parents of L2 refined at L1).

With ``BatchConfig.autotune`` every kernel call of the first regrid(s) and a
tuning pass over the step kernels run under the schedule autotuner
(``subsetix_amd.autotune``): each shape bucket times its candidate schedules
(segment tiles or row groups of 4 / 16 / 64 lanes, fused or two-pass set
kernels) and the captured steps use the fastest; later regrids take the
nearest tuned bucket's choice.
"""

from __future__ import annotations

import contextlib
import math
import time
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _core, _lib, autotune
from ._lib import _I32, _I64, _P, check, context, ptr
from .expressions import IntervalSet, binary_op
from .interval_field import IntervalField, cell_offsets, create_interval_field, known_cell_count, remember_cell_count
from .interval_stencil import build_stencil_plan
from .levels import gradient_tag, prolong_into, restrict_into, transfer_plan, upwind_batched
from .morphology import dilate_interval_set
from .multilevel import prolong_set, restrict_set



@dataclass
class BatchConfig:
    levels: int = 3
    ratio: int = 2
    cfl: float = 0.5
    threshold: tuple = (0.02, 0.02)  # gradient threshold per tagged level (per cell)
    grading: int = 1
    gap: int = 16
    regrid_every: int = 8
    use_graph: bool = True
    accumulate: str = "f64"
    autotune: bool = False  # time the candidate schedules of every kernel per shape bucket (regrid + step)
    autotune_regrids: int = 1  # regrids that tune; later ones reuse the nearest tuned buckets


class BatchedAMR:
    def __init__(self, instances: List[dict], *, dtype=torch.float32, config: Optional[BatchConfig] = None,
                 device=None):
        self.cfg = config or BatchConfig()
        if self.cfg.levels < 1 or self.cfg.levels > 3:
            raise ValueError("levels must be 1, 2 or 3")
        self.dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.dtype = dtype
        self.instances = instances
        self.n = len(instances)
        R = self.cfg.ratio
        sizes = np.array([int(i["n"]) for i in instances], dtype=np.int64)
        self.sizes = sizes
        gap = self.cfg.gap
        base0 = np.zeros(self.n, np.int64)
        base0[1:] = np.cumsum(sizes + gap)[:-1]
        top = int((base0[-1] + sizes[-1] + gap) * R ** (self.cfg.levels - 1))
        if top >= 2 ** 31 - 1:
            raise ValueError("batch too large for int32 row ids")
        self.bases = [torch.from_numpy((base0 * R ** L).astype(np.int32)).to(self.dev) for L in range(self.cfg.levels)]
        self.domains = [self._domain(base0 * R ** L, sizes * R ** L) for L in range(self.cfg.levels)]
        # per-level instance tables
        vel = np.array([i["velocity"] for i in instances], dtype=np.float64)
        finest = sizes * R ** (self.cfg.levels - 1)
        denom = np.abs(vel[:, 0]) * finest + np.abs(vel[:, 1]) * finest
        dts = np.where(denom > 0, self.cfg.cfl / np.maximum(denom, 1e-300), 0.0)
        self.dts = dts
        npdt = np.float32 if dtype == torch.float32 else np.float64
        self.inst_w, self.inst_coef = [], []
        for L in range(self.cfg.levels):
            nL = sizes * R ** L
            coef = np.stack([vel[:, 0], vel[:, 1], dts, 1.0 / nL, 1.0 / nL], axis=1).astype(npdt).reshape(-1)
            self.inst_w.append(torch.from_numpy(nL.astype(np.int32)).to(self.dev))
            self.inst_coef.append(torch.from_numpy(coef).to(self.dev))
        # level-0 field
        f0 = create_interval_field(self.domains[0], 0.0, dtype=dtype)
        cs = np.zeros(self.n + 1, np.int64)
        cs[1:] = np.cumsum(sizes * sizes)
        self.cstart0 = torch.from_numpy(cs).to(self.dev)
        self.cell_area0 = torch.from_numpy((1.0 / (sizes * sizes)).astype(np.float64)).to(self.dev)
        self._init_values(f0)
        self.f = [f0]
        self.buf = [[f0.values, torch.empty_like(f0.values)]]
        self.cur = 0
        self.graph = None
        self.step_count = 0
        self.timings: Dict[str, float] = {}
        self._regrids = 0
        self._plan0 = None
        self.regrid()
        self._last_regrid = 0

    # ------------------------------------------------------------------ setup
    def _domain(self, base: np.ndarray, n: np.ndarray) -> IntervalSet:
        rows = np.concatenate([np.arange(b, b + k, dtype=np.int64) for b, k in zip(base, n)]).astype(np.int32)
        widths = np.repeat(n, n).astype(np.int32)
        off = np.arange(rows.size + 1, dtype=np.int32)
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(self.dev)
        return IntervalSet._trusted(t(np.zeros(rows.size, np.int32)), t(widths), t(off), t(rows), dense=False)

    def _init_values(self, f0: IntervalField) -> None:
        off = 0
        for inst, n in zip(self.instances, self.sizes):
            n = int(n)
            xs = (torch.arange(n, device=self.dev, dtype=torch.float64) + 0.5) / n
            Y, X = torch.meshgrid(xs, xs, indexing="ij")
            u = torch.zeros((n, n), dtype=torch.float64, device=self.dev)
            for cx, cy, sg, amp in inst["bumps"]:
                u += amp * torch.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sg * sg))
            f0.values[off:off + n * n].copy_(u.reshape(-1).to(self.dtype))
            off += n * n

    def _row_inst(self, s: IntervalSet, level: int) -> torch.Tensor:
        out = _core.empty_i32(s.row_count, self.dev)
        ctx = context(self.dev)
        check(_lib.lib().stx_row_instance(ptr(s.rows), s.row_count, ptr(self.bases[level]), self.n, ptr(out),
                                          ctx.stream), "row_instance")
        return out

    # ------------------------------------------------------------------ regrid
    def _tag(self, fld: IntervalField, level: int, plan=None) -> IntervalSet:
        # per-cell threshold scaled to the level's resolution: |grad u| per
        # cell halves when the cell size halves
        t = self.cfg.threshold[min(level, len(self.cfg.threshold) - 1)]
        return gradient_tag(fld, t, plan=plan)

    def regrid(self) -> None:
        t0 = time.perf_counter()
        tune = self.cfg.autotune and self._regrids < self.cfg.autotune_regrids
        self._regrids += 1
        with autotune.tuning() if tune else contextlib.nullcontext():
            self._regrid()
            if tune:
                # tuning pass over the step kernels on the new geometry, before
                # any capture: it leaves the state as it was (the ghost
                # prolongation is idempotent, the other kernels write only the
                # next buffers, which a step overwrites whole)
                self._step_kernels(self.cur)
        torch.cuda.synchronize(self.dev)
        self.timings["regrid_s"] = self.timings.get("regrid_s", 0.0) + time.perf_counter() - t0

    def _regrid(self) -> None:
        R = self.cfg.ratio
        p = self.cfg.grading
        L = self.cfg.levels
        D = self.domains
        cur0 = IntervalField(D[0], self.buf[0][self.cur], self.f[0].interval_cell_offsets)
        old = self._fields(self.cur) if hasattr(self, "ext") else list(self.f)
        old_sets = list(getattr(self, "level_sets", []))
        if self._plan0 is None:
            # level 0 is the fixed full domain of every instance: one segment
            # table serves its gradient tags and its stencil for the whole run
            self._plan0 = build_stencil_plan(D[0], self.f[0].interval_cell_offsets, 0, self._row_inst(D[0], 0),
                                             self.inst_w[0])
        tag0 = self._tag(cur0, 0, self._plan0)
        refine0 = binary_op(dilate_interval_set(tag0, p, p) if p > 0 else tag0, D[0], "intersection")
        sets = [D[0]]
        refines = []
        if L >= 2:
            ring = binary_op(dilate_interval_set(prolong_set(refine0, R), 1, 1), D[1], "intersection")
            refine0 = binary_op(refine0, restrict_set(ring, R), "union")
            l1 = prolong_set(refine0, R)
            sets.append(l1)
            refines.append(refine0)
        if L >= 3:
            # level-1 values for tagging: fresh prolongation, old values kept
            tmp1 = self._level_field(1, l1, cur0, old, old_sets, ghost=None)
            comp = binary_op(D[1], l1, "difference")
            eroded = binary_op(l1, dilate_interval_set(comp, 1, 1), "difference")
            # a segment table for the tag (every level-1 cell lies inside its
            # instance's width): bit-mask passes instead of row walks
            plan1 = build_stencil_plan(l1, tmp1.interval_cell_offsets, 0, self._row_inst(l1, 1), self.inst_w[1])
            tag1 = self._tag(tmp1, 1, plan1)
            refine1 = binary_op(dilate_interval_set(tag1, p, p) if p > 0 else tag1, eroded, "intersection")
            l2 = prolong_set(refine1, R)
            sets.append(l2)
            refines.append(refine1)
        # ghosts and extended level fields
        self.level_sets = sets
        self.refines = refines
        self.ghosts = [None]
        self.ext = [D[0]]
        new_f = [cur0]
        for lv in range(1, L):
            g = binary_op(binary_op(dilate_interval_set(sets[lv], 1, 1), D[lv], "intersection"), sets[lv],
                          "difference")
            ext = binary_op(sets[lv], g, "union")
            self.ghosts.append(g)
            self.ext.append(ext)
            new_f.append(self._level_field(lv, ext, new_f[lv - 1], old, old_sets, ghost=g))
        self.f = new_f
        self.buf = [[new_f[0].values, self.buf[0][1 - self.cur]]] + [[fl.values, torch.empty_like(fl.values)]
                                                                    for fl in new_f[1:]]
        self.cur = 0
        self.row_inst = [self._row_inst(self.ext[lv], lv) for lv in range(L)]
        for t in self.ghosts[1:] + self.refines:
            cell_offsets(t)
        # transfer and stencil tables, built before any graph capture
        if R == 2:
            fl = self._fields(self.cur)
            for lv in range(1, L):
                if self.ghosts[lv].interval_count:
                    transfer_plan(0, self.ghosts[lv], fl[lv], fl[lv - 1])
                if self.refines[lv - 1].interval_count:
                    transfer_plan(1, self.refines[lv - 1], fl[lv - 1], fl[lv])
        self.plans = [self._plan0] + [build_stencil_plan(self.ext[lv], self.f[lv].interval_cell_offsets, 0,
                                                         self.row_inst[lv], self.inst_w[lv]) for lv in range(1, L)]
        self.graph = None
        self.cells = self.cells_per_step()

    def _level_field(self, lv, target, coarser: IntervalField, old, old_sets, ghost) -> IntervalField:
        """Field on `target` at level lv: prolonged from the coarser level,
        then the previous values copied back where the previous level lv
        existed (not its ghosts)."""
        f = create_interval_field(target, 0.0, dtype=self.dtype)
        prolong_into(target, f, coarser, self.cfg.ratio)
        if lv < len(old) and lv < len(old_sets):
            prev = IntervalField(old[lv].interval_set, old[lv].values, old[lv].interval_cell_offsets)
            keep = binary_op(target, old_sets[lv], "intersection")
            prolong_into(keep, f, prev, 1)
        return f

    # ------------------------------------------------------------------ step
    def _fields(self, idx: int) -> List[IntervalField]:
        return [IntervalField(self.ext[lv], self.buf[lv][idx], self.f[lv].interval_cell_offsets)
                for lv in range(self.cfg.levels)]

    def _step_kernels(self, idx: int) -> None:
        R = self.cfg.ratio
        L = self.cfg.levels
        cur = self._fields(idx)
        nxt = self._fields(1 - idx)
        for lv in range(1, L):
            prolong_into(self.ghosts[lv], cur[lv], cur[lv - 1], R)
        for lv in range(L):
            upwind_batched(cur[lv], cur[lv].values, nxt[lv].values, self.row_inst[lv], self.inst_w[lv],
                           self.inst_coef[lv], plan=self.plans[lv])
        for lv in range(L - 1, 0, -1):
            restrict_into(self.refines[lv - 1], nxt[lv - 1], nxt[lv], R, accumulate=self.cfg.accumulate)

    def _capture(self, start: int) -> "torch.cuda.CUDAGraph":
        """Graph of two steps starting from buffer `start` (ends on it)."""
        if not hasattr(self, "_gstream"):
            self._gstream = torch.cuda.Stream(self.dev)
        self._gstream.wait_stream(torch.cuda.current_stream(self.dev))
        with torch.cuda.stream(self._gstream):
            context(self.dev)  # per-stream context allocated outside the capture
        # capture_begin / capture_end on the side stream: the graph context
        # manager would synchronise the device and empty the allocator cache
        # at every capture, and a capture follows every regrid
        g = torch.cuda.CUDAGraph()
        with torch.cuda.stream(self._gstream):
            g.capture_begin()
            try:
                self._step_kernels(start)
                self._step_kernels(1 - start)
            finally:
                g.capture_end()
        torch.cuda.current_stream(self.dev).wait_stream(self._gstream)
        return g

    def _advance(self, n: int) -> None:
        if self.cfg.use_graph and n >= 2:
            if self.graph is None:
                self.graph = {}
            g = self.graph.get(self.cur)
            if g is None:
                g = self._capture(self.cur)
                self.graph[self.cur] = g
            for _ in range(n // 2):
                g.replay()
            n -= 2 * (n // 2)
        for _ in range(n):
            self._step_kernels(self.cur)
            self.cur = 1 - self.cur

    def run(self, steps: int) -> None:
        """Advance every instance by `steps` steps, regridding every k steps."""
        k = self.cfg.regrid_every
        remaining = int(steps)
        while remaining > 0:
            if self.step_count % k == 0 and getattr(self, "_last_regrid", 0) != self.step_count:
                self.regrid()
                self._last_regrid = self.step_count
            n = min(k - self.step_count % k, remaining)
            self._advance(n)
            remaining -= n
            self.step_count += n

    # ------------------------------------------------------------------ stats
    def cells_per_step(self) -> Dict[str, int]:
        """Cells of every level set and of the ghost rings (one host read)."""
        sets = list(self.level_sets) + list(self.ghosts[1:])
        known = [known_cell_count(s, cell_offsets(s)) for s in sets]
        ends = [cell_offsets(s)[-1:] for s, k in zip(sets, known) if k is None]
        vals = iter(torch.cat(ends).tolist() if ends else [])
        cells = [int(next(vals)) if k is None else k for k in known]
        for s, c in zip(sets, cells):
            remember_cell_count(s, cell_offsets(s), c)
        lv = {f"level{i}": cells[i] for i in range(self.cfg.levels)}
        lv["ghost"] = sum(cells[self.cfg.levels:])
        return lv

    def masses(self) -> torch.Tensor:
        """FP64 mass of every instance (level 0 after the restrictions)."""
        out = torch.empty(self.n, dtype=torch.float64, device=self.dev)
        ctx = context(self.dev)
        v = self.buf[0][self.cur]
        check(_lib.lib().stx_instance_mass(4 if self.dtype == torch.float32 else 8, ptr(v), ptr(self.cstart0), self.n,
                                           ptr(self.cell_area0), ptr(out), ctx.stream), "instance_mass")
        return out

    def level0_values(self) -> torch.Tensor:
        return self.buf[0][self.cur]
