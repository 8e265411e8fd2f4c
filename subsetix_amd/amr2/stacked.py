"""Several two-level composites advanced as one: the instances of
``CompositeTwoLevel`` (full coarse level, fine level on the refined region
plus a one-cell ghost ring, ratio 2) stacked row-wise with empty rows between
them, as in ``subsetix_amd.batch``: one interval set per level holds every
instance, so building the geometry is one chain of set operations for the
whole stack and a step is four launches (ghost prolongation, upwind on each
level with per-instance coefficients, restriction with FP64 sums) whatever
the number of instances.

Every kernel is row- or cell-local and instances never share a row, so each
instance's fields are bit-identical to a ``CompositeTwoLevel`` run of that
instance alone (same initial values, same coefficients rounded the same way);
``snapshot`` copies one instance's fields out at the step count it needs.
"""

from __future__ import annotations

from typing import Dict, List, Optional

import numpy as np
import torch

from .. import _core, _lib
from .._lib import check, context, ptr
from ..expressions import IntervalSet, binary_op
from ..interval_field import IntervalField, cell_offsets, create_interval_field
from ..interval_stencil import build_stencil_plan
from ..levels import prolong_into, restrict_into, transfer_plan, upwind_batched
from ..morphology import dilate_interval_set
from ..multilevel import prolong_set, restrict_set
from .mass import total_mass_async


def _stack_domain(base: np.ndarray, n: np.ndarray, dev) -> IntervalSet:
    """Full n_i x n_i domains at row offsets base_i (one interval per row)."""
    rows = np.concatenate([np.arange(b, b + k, dtype=np.int64) for b, k in zip(base, n)]).astype(np.int32)
    widths = np.repeat(n, n).astype(np.int32)
    off = np.arange(rows.size + 1, dtype=np.int32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return IntervalSet._trusted(t(np.zeros(rows.size, np.int32)), t(widths), t(off), t(rows), dense=False)


def _stack_host_sets(sets: List[dict], shifts: np.ndarray, dev) -> IntervalSet:
    """Host CSR sets (rows, off, begin, end) with row ids shifted, as one set."""
    rows = np.concatenate([h["rows"].astype(np.int64) + s for h, s in zip(sets, shifts)]).astype(np.int32)
    beg = np.concatenate([h["begin"] for h in sets]).astype(np.int32)
    end = np.concatenate([h["end"] for h in sets]).astype(np.int32)
    offs, base = [np.zeros(1, np.int64)], 0
    for h in sets:
        offs.append(h["off"][1:].astype(np.int64) + base)
        base += int(h["off"][-1])
    off = np.concatenate(offs).astype(np.int32)
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return IntervalSet._trusted(t(beg), t(end), t(off), t(rows), dense=None)


class StackedComposite:
    """``instances``: dicts with ``discs`` (host CSR of the refined fine
    cells: rows, off, begin, end), ``n_fine`` (even) and ``velocity``."""

    def __init__(self, instances: List[dict], *, dtype=torch.float32, device=None, cfl: float = 0.5,
                 accumulate: str = "f64", gap: int = 16):
        self.dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.dtype = dtype
        self.accumulate = accumulate
        self.n = len(instances)
        if self.n == 0:
            raise ValueError("no instances")
        nf = np.array([int(i["n_fine"]) for i in instances], dtype=np.int64)
        if np.any(nf % 2) or np.any(nf < 2):
            raise ValueError("n_fine must be even and positive")
        nc = nf // 2
        self.nc, self.nf = nc, nf
        base = np.zeros(self.n, np.int64)
        base[1:] = np.cumsum(nc + gap)[:-1]
        if int((base[-1] + nc[-1] + gap) * 2) >= 2 ** 31 - 1:
            raise ValueError("stack too large for int32 row ids")
        self.base = base
        dev = self.dev
        # geometry of the whole stack (one chain of set operations)
        self.coarse_set = _stack_domain(base, nc, dev)
        fine_domain = _stack_domain(2 * base, nf, dev)
        discs = _stack_host_sets([i["discs"] for i in instances], 2 * base, dev)
        self.refine = binary_op(restrict_set(discs, 2), self.coarse_set, "intersection")
        self.fine = prolong_set(self.refine, 2)
        dil = binary_op(dilate_interval_set(self.fine, 1, 1), fine_domain, "intersection")
        self.ghost = binary_op(dil, self.fine, "difference")
        self.ext = binary_op(self.fine, self.ghost, "union")
        c0 = create_interval_field(self.coarse_set, 0.0, dtype=dtype)
        f0 = create_interval_field(self.ext, 0.0, dtype=dtype)
        self._c_co, self._f_co = c0.interval_cell_offsets, f0.interval_cell_offsets
        self._cbuf = [c0.values, torch.empty_like(c0.values)]
        self._fbuf = [f0.values, torch.empty_like(f0.values)]
        self.cur = 0
        # per-instance cell ranges: coarse levels are full domains in order;
        # the extended fine level's instance j owns rows [2 base_j, 2 base_j + nf_j)
        cs = np.zeros(self.n + 1, np.int64)
        cs[1:] = np.cumsum(nc * nc)
        self.cstart = cs
        rows = self.ext.rows.cpu().numpy().astype(np.int64)
        off = self.ext.row_offsets.cpu().numpy().astype(np.int64)
        co = self._f_co.cpu().numpy().astype(np.int64)
        k0 = np.searchsorted(rows, 2 * base)
        k1 = np.searchsorted(rows, 2 * base + nf)
        self.fstart = co[off[k0]]
        self.fend = co[off[k1]]
        self.fine_cells = np.zeros(self.n, np.int64)  # fine cells without ghosts, per instance
        frows = self.fine.rows.cpu().numpy().astype(np.int64)
        foff = self.fine.row_offsets.cpu().numpy().astype(np.int64)
        fco = cell_offsets(self.fine).cpu().numpy().astype(np.int64) if self.fine.interval_count else np.zeros(1)
        j0 = np.searchsorted(frows, 2 * base)
        j1 = np.searchsorted(frows, 2 * base + nf)
        if self.fine.interval_count:
            self.fine_cells = fco[foff[j1]] - fco[foff[j0]]
        self.ghost_cells = (self.fend - self.fstart) - self.fine_cells
        # per-instance coefficients, computed as CompositeTwoLevel computes them
        npdt = np.float32 if dtype == torch.float32 else np.float64
        coef_c, coef_f = [], []
        self.dt = np.zeros(self.n)
        for j, inst in enumerate(instances):
            a, b = float(inst["velocity"][0]), float(inst["velocity"][1])
            dx_c = 1.0 / int(nc[j])
            dy_c = 1.0 / int(nc[j])
            dx_f = dx_c / 2
            dy_f = dy_c / 2
            denom = abs(a) / dx_f + abs(b) / dy_f
            dt = cfl / denom if denom > 0 else 0.0
            self.dt[j] = dt
            coef_c += [a, b, dt, dx_c, dy_c]
            coef_f += [a, b, dt, dx_f, dy_f]
        # as CompositeTwoLevel.mass(): dx_c * dy_c
        self.cell_area = np.array([(1.0 / int(k)) * (1.0 / int(k)) for k in nc])
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        self.inst_coef = [t(np.array(coef_c, dtype=npdt)), t(np.array(coef_f, dtype=npdt))]
        self.inst_w = [t(nc.astype(np.int32)), t(nf.astype(np.int32))]
        bases = [t(base.astype(np.int32)), t((2 * base).astype(np.int32))]
        self.row_inst = [self._row_inst(self.coarse_set, bases[0]), self._row_inst(self.ext, bases[1])]
        self.plans = [build_stencil_plan(self.coarse_set, self._c_co, 0, self.row_inst[0], self.inst_w[0]),
                      build_stencil_plan(self.ext, self._f_co, 0, self.row_inst[1], self.inst_w[1])]
        for s in (self.ghost, self.refine):
            if s.interval_count:
                cell_offsets(s)
        c, f = self._fields(0)
        if self.ghost.interval_count:
            transfer_plan(0, self.ghost, f, c)
        if self.refine.interval_count:
            transfer_plan(1, self.refine, c, f)
        self._graphs: Dict[tuple, torch.cuda.CUDAGraph] = {}

    def _row_inst(self, s: IntervalSet, bases: torch.Tensor) -> torch.Tensor:
        out = _core.empty_i32(s.row_count, self.dev)
        ctx = context(self.dev)
        check(_lib.lib().stx_row_instance(ptr(s.rows), s.row_count, ptr(bases), self.n, ptr(out), ctx.stream),
              "row_instance")
        return out

    def _fields(self, idx: int):
        return (IntervalField(self.coarse_set, self._cbuf[idx], self._c_co),
                IntervalField(self.ext, self._fbuf[idx], self._f_co))

    # ----------------------------------------------------------------- values
    def set_instance_values(self, j: int, coarse_values: torch.Tensor, fine_values: torch.Tensor) -> None:
        """Initial values of instance j (its coarse cells row-major, its
        extended fine cells in storage order); call ``finish_init`` after the
        last instance."""
        c0, c1 = int(self.cstart[j]), int(self.cstart[j + 1])
        f0, f1 = int(self.fstart[j]), int(self.fend[j])
        self._cbuf[self.cur][c0:c1].copy_(coarse_values.to(self.dtype))
        self._fbuf[self.cur][f0:f1].copy_(fine_values.to(self.dtype))

    def finish_init(self) -> None:
        """Restriction onto the refined coarse cells and the ghost
        prolongation, as CompositeTwoLevel.set_values."""
        c, f = self._fields(self.cur)
        restrict_into(self.refine, c, f, 2, accumulate=self.accumulate)
        prolong_into(self.ghost, f, c, 2)

    def ext_cells(self, j: int) -> int:
        return int(self.fend[j] - self.fstart[j])

    # ------------------------------------------------------------------ steps
    def _kernels(self, cur: int) -> None:
        c, f = self._fields(cur)
        nc_, nf_ = self._fields(1 - cur)
        prolong_into(self.ghost, f, c, 2)
        upwind_batched(c, c.values, nc_.values, self.row_inst[0], self.inst_w[0], self.inst_coef[0],
                       plan=self.plans[0])
        upwind_batched(f, f.values, nf_.values, self.row_inst[1], self.inst_w[1], self.inst_coef[1],
                       plan=self.plans[1])
        restrict_into(self.refine, nc_, nf_, 2, accumulate=self.accumulate)

    def step(self) -> None:
        self._kernels(self.cur)
        self.cur = 1 - self.cur

    def _graph(self, start: int, nsteps: int, stream: torch.cuda.Stream) -> torch.cuda.CUDAGraph:
        """Graph of ``nsteps`` steps from buffer ``start``, captured on
        ``stream`` with capture_begin / capture_end (the graph context
        manager would synchronise the device and empty the allocator cache on
        every capture, stalling the other streams' hosts)."""
        key = (start, nsteps)
        g = self._graphs.get(key)
        if g is None:
            with torch.cuda.stream(stream):
                context(self.dev)  # per-stream context allocated outside the capture
                g = torch.cuda.CUDAGraph()
                g.capture_begin()
                try:
                    cur = start
                    for _ in range(nsteps):
                        self._kernels(cur)
                        cur = 1 - cur
                finally:
                    g.capture_end()
            self._graphs[key] = g
        return g

    def advance(self, n: int, stream: Optional[torch.cuda.Stream] = None, use_graph: bool = True) -> None:
        """``n`` steps on ``stream`` (the current stream when None): eager
        launches, or with ``use_graph`` replays of captured two-step and
        one-step graphs (no host synchronisation either way)."""
        if stream is None:
            for _ in range(n):
                self.step()
            return
        if not use_graph:
            with torch.cuda.stream(stream):
                for _ in range(n):
                    self.step()
            return
        with torch.cuda.stream(stream):
            while n >= 2:
                self._graph(self.cur, 2, stream).replay()
                n -= 2
            if n:
                self._graph(self.cur, 1, stream).replay()
                self.cur = 1 - self.cur

    # ---------------------------------------------------------------- results
    def snapshot(self, j: int) -> dict:
        """Copies of instance j's current coarse and extended fine values and
        its FP64 mass (device tensors, no host synchronisation)."""
        c0, c1 = int(self.cstart[j]), int(self.cstart[j + 1])
        f0, f1 = int(self.fstart[j]), int(self.fend[j])
        coarse = self._cbuf[self.cur][c0:c1].clone()
        fine = self._fbuf[self.cur][f0:f1].clone()
        return {"coarse": coarse, "fine": fine, "mass": total_mass_async(coarse, float(self.cell_area[j]))}

    def retire(self, j: int) -> None:
        """Stop updating instance j (its upwind is skipped from the next
        launch on; call after its last snapshot).  A device write on the
        current stream, outside any captured graph, which reads the
        coefficients at every replay."""
        for coef in self.inst_coef:
            coef[5 * j + 3:5 * j + 4].zero_()  # a fill kernel, no host synchronisation

    def cells_per_step(self, j: int) -> int:
        """Cell updates of instance j per step (coarse + fine + ghost)."""
        return int(self.cstart[j + 1] - self.cstart[j]) + int(self.fend[j] - self.fstart[j])
