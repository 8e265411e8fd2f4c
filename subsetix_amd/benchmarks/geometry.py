"""Deterministic geometries and fields of the benchmark suite (the inputs of
subsetix_cupy/benchmarks/utils.py, built here with vectorised NumPy run
extraction on the host, then one copy to the device).

* ``square_grid_interval_set``: a tiles_x x tiles_y grid of filled squares
  (fill_ratio of each tile, centred), optionally rolled by (shift_x,
  shift_y) with wrap-around, as maximal runs per row;
* ``staircase_interval_set``: the rows of a reference set shifted right in
  steps, keeping every row's interval count and lengths;
* ``random_interval_set``: per row, a seeded number of disjoint intervals
  from sorted distinct end points (the same generator calls, in the same
  order, as the reference, so the sets are identical);
* ``deterministic_interval_field``: u = amplitude * sin(2 pi f x / W +
  phase + pi * row / R) on every cell.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Tuple

import numpy as np
import torch

from ..expressions import IntervalSet, build_interval_set
from ..interval_field import IntervalField, create_interval_field


@dataclass(frozen=True)
class GeometrySpec:
    rows: int
    width: int
    tiles_x: int
    tiles_y: int
    fill_ratio: float = 0.75


def _runs_of_mask(mask: np.ndarray):
    """(row_offsets, begin, end) of the maximal runs of every row of a
    boolean mask."""
    h, w = mask.shape
    pad = np.zeros((h, w + 2), np.int8)
    pad[:, 1:-1] = mask
    d = np.diff(pad, axis=1)
    rs, bs = np.nonzero(d == 1)
    _, es = np.nonzero(d == -1)
    off = np.zeros(h + 1, np.int64)
    np.cumsum(np.bincount(rs, minlength=h), out=off[1:])
    return off.astype(np.int32), bs.astype(np.int32), es.astype(np.int32)


def square_grid_mask(spec: GeometrySpec, shift_x: int = 0, shift_y: int = 0) -> np.ndarray:
    rows, width = spec.rows, spec.width
    tx, ty = max(1, min(4, spec.tiles_x)), max(1, spec.tiles_y)
    fill = float(np.clip(spec.fill_ratio, 0.05, 1.0))
    ch, cw = rows / ty, width / tx
    sh, sw = max(1, int(round(ch * fill))), max(1, int(round(cw * fill)))
    ys = np.zeros(rows, bool)
    xs = np.zeros(width, bool)
    for gy in range(ty):
        yc = (gy + 0.5) * ch
        y0, y1 = (max(0, min(rows, int(round(v)))) for v in (yc - sh / 2, yc + sh / 2))
        ys[y0:y1] = True
    for gx in range(tx):
        xc = (gx + 0.5) * cw
        x0, x1 = (max(0, min(width, int(round(v)))) for v in (xc - sw / 2, xc + sw / 2))
        xs[x0:x1] = True
    mask = np.outer(ys, xs)
    if shift_y:
        mask = np.roll(mask, int(shift_y), axis=0)
    if shift_x:
        mask = np.roll(mask, int(shift_x), axis=1)
    return mask


def square_grid_csr(spec: GeometrySpec, shift_x: int = 0, shift_y: int = 0):
    """Host (row_offsets, begin, end) of the square grid."""
    return _runs_of_mask(square_grid_mask(spec, shift_x, shift_y))


def square_grid_interval_set(*, spec: GeometrySpec, shift_x: int = 0, shift_y: int = 0, device=None) -> IntervalSet:
    off, b, e = square_grid_csr(spec, shift_x, shift_y)
    return build_interval_set(off, b, e, device=device)


def staircase_csr(width: int, off, beg, end, steps: int = 8):
    """Host (row_offsets, begin, end) of the staircase built from a
    reference CSR (rows keep their interval counts and lengths)."""
    if steps <= 0:
        raise ValueError("steps has to be at least 1")
    off = np.asarray(off, np.int64)
    beg = np.asarray(beg, np.int64)
    end = np.asarray(end, np.int64)
    counts = np.diff(off)
    active = np.nonzero(counts > 0)[0]
    if active.size == 0:
        return off.astype(np.int32), beg.astype(np.int32), end.astype(np.int32)
    first_b = beg[off[active]]
    span = end[off[active + 1] - 1] - first_b
    avail = max(0, width - int(span.max()))
    nsteps = min(steps, max(1, avail + 1))
    step_h = max(1, active.size // nsteps)
    stride = max(1, avail // max(1, nsteps - 1)) if nsteps > 1 else 0
    shift = np.minimum((np.arange(active.size) // step_h) * stride, np.maximum(0, width - span))
    # every interval moves with its row: shift + (x - row's first begin)
    pos_of_row = np.full(off.size - 1, -1, np.int64)
    pos_of_row[active] = np.arange(active.size)
    p = pos_of_row[np.repeat(np.arange(off.size - 1), counts)]
    nb = shift[p] + (beg - first_b[p])
    ne = shift[p] + (end - first_b[p])
    over = np.maximum(ne - width, 0)
    return off.astype(np.int32), np.maximum(nb - over, 0).astype(np.int32), (ne - over).astype(np.int32)


def staircase_interval_set(*, spec: GeometrySpec, reference: IntervalSet | None = None, steps: int = 8,
                           device=None) -> IntervalSet:
    base = reference if reference is not None else square_grid_interval_set(spec=spec, device=device)
    if base.interval_count == 0:
        return base
    off, b, e = staircase_csr(spec.width, base.row_offsets.cpu().numpy(), base.begin.cpu().numpy(),
                              base.end.cpu().numpy(), steps)
    return build_interval_set(off, b, e, device=base.device)


def random_csr(*, rows: int, width: int, intervals_per_row: Tuple[int, int] | int, seed: int = 0):
    """Host (row_offsets, begin, end) of the seeded random set."""
    rng = np.random.default_rng(seed)
    if isinstance(intervals_per_row, int):
        lo = hi = max(1, intervals_per_row)
    else:
        lo = max(1, intervals_per_row[0])
        hi = max(lo, intervals_per_row[1])
    points = np.arange(0, width + 1, dtype=np.int32)
    off, bs, es = [0], [], []
    for _ in range(rows):
        k = int(rng.integers(lo, hi + 1))
        k = 0 if width <= 1 else min(k, max(0, width // 2))
        if k > 0:
            p = np.sort(rng.choice(points, size=2 * k, replace=False))
            b, e = p[0::2].astype(np.int64), np.clip(np.maximum(p[1::2], p[0::2] + 1), 0, width).astype(np.int64)
            last = 0
            for s, t in zip(b.tolist(), e.tolist()):
                s = max(s, last)
                if s >= width:
                    break
                t = min(max(t, s + 1), width)
                if t <= s:
                    continue
                bs.append(s)
                es.append(t)
                last = t
        off.append(len(bs))
    return np.array(off, np.int32), np.array(bs, np.int32), np.array(es, np.int32)


def random_interval_set(*, rows: int, width: int, intervals_per_row: Tuple[int, int] | int, seed: int = 0,
                        device=None) -> IntervalSet:
    off, b, e = random_csr(rows=rows, width=width, intervals_per_row=intervals_per_row, seed=seed)
    return build_interval_set(off, b, e, device=device)


def deterministic_interval_field(interval_set: IntervalSet, *, width: int, amplitude: float = 1.0,
                                 frequency: float = 0.125, phase: float = 0.0) -> IntervalField:
    field = create_interval_field(interval_set, fill_value=0.0, dtype=torch.float32)
    n = field.values.numel()
    if n == 0:
        return field
    off = interval_set.row_offsets.cpu().numpy().astype(np.int64)
    beg = interval_set.begin.cpu().numpy().astype(np.int64)
    end = interval_set.end.cpu().numpy().astype(np.int64)
    lens = np.maximum(end - beg, 0)
    row_ord = np.repeat(np.arange(off.size - 1), np.diff(off))
    start = np.concatenate([[0], np.cumsum(lens)[:-1]])
    x = np.arange(lens.sum(), dtype=np.int64) + np.repeat(beg - start, lens)
    rphase = phase + np.repeat(row_ord, lens) * (1.0 / max(1, interval_set.row_count)) * np.pi
    vals = amplitude * np.sin(2.0 * np.pi * frequency * x * (1.0 / max(1, width)) + rphase)
    field.values.copy_(torch.from_numpy(vals.astype(np.float32)))
    return field


def random_interval_field(interval_set: IntervalSet, *, seed: int = 0, dtype=torch.float32) -> IntervalField:
    field = create_interval_field(interval_set, fill_value=0.0, dtype=dtype)
    if field.values.numel():
        rng = np.random.default_rng(seed)
        field.values.copy_(torch.from_numpy(rng.standard_normal(field.values.numel()).astype(np.float32)))
    return field


__all__ = ["GeometrySpec", "square_grid_mask", "square_grid_csr", "square_grid_interval_set", "staircase_csr",
           "staircase_interval_set", "random_csr", "random_interval_set", "deterministic_interval_field",
           "random_interval_field"]
