"""Level-to-level kernels on interval fields: fused transfers onto a target
subset, gradient tagging on sparse levels and the batched upwind step.

These are the building blocks of the sparse multi-level drivers
(``subsetix_amd.amr2.composite`` and ``subsetix_amd.batch``): unlike the
reference's synchronisation (restrict the whole fine field, gather, scatter;
fields.py:570-612) each transfer is one pass over the target cells.
"""

from __future__ import annotations

import torch

from . import _core, _lib
from ._lib import _F64, _I32, _I64, _P, check, context, ptr
from .expressions import IntervalSet
from .interval_field import IntervalField, cell_offsets

_lib.register("stx_transfer", _I32, [_I32, _I32, _I32, _I32, _I32, _P, _P, _P, _P, _P, _I64, _P, _P, _P, _P, _P, _I64,
                                     _P, _P, _P, _P, _P, _I64, _P, _P, _P])
_lib.register("stx_grad_tag", _I32, [_I32, _I32, _P, _P, _I64, _P, _P, _P, _P, _F64, _P, _P, _P, _I64, _P, _P, _P,
                                     _I64, _I32, _P])
_lib.register("stx_upwind_batched", _I32, [_I32, _I32, _P, _P, _I64, _P, _P, _P, _P, _P, _P, _P, _P, _P])
_GEOM7 = [_P, _P, _P, _P, _P, _I64]
_lib.register("stx_transfer_tile_cells", _I64, [])
_lib.register("stx_transfer_seg_count", _I32, [_I32] + _GEOM7 + [_P, _I64] + _GEOM7 + _GEOM7
              + [_P, _P, _P, _P, _P, _I64, _I32, _P])
_lib.register("stx_transfer_seg_build", _I32, [_I32] + _GEOM7 + [_P, _I64] + _GEOM7 + _GEOM7
              + [_P, _I64, _P, _I64, _P, _I64, _P])
_lib.register("stx_transfer_seg", _I32, [_I32, _I32, _I32, _P, _I64, _I64, _P, _P, _I64, _P, _P])


def _geom(s: IntervalSet, co: torch.Tensor):
    return (ptr(s.rows), ptr(s.row_offsets), ptr(s.begin), ptr(s.end), ptr(co), s.row_count)


def _group(cells: int, rows: int) -> int:
    avg = cells / max(rows, 1)
    return 64 if avg >= 96 else 16 if avg >= 12 else 4


def _dtype_code(t: torch.Tensor) -> int:
    if t.dtype == torch.float32:
        return 4
    if t.dtype == torch.float64:
        return 8
    raise TypeError("transfers need float32 or float64 fields")


def transfer_plan(mode: int, target: IntervalSet, dst: IntervalField, src: IntervalField):
    """Segment table of a ratio-2 transfer onto ``target`` (cached on the
    target set, keyed by the destination and source geometries)."""
    cache = target._meta.setdefault("transfer_plans", {})
    key = (mode, id(dst.interval_set), id(src.interval_set), id(dst.interval_cell_offsets),
           id(src.interval_cell_offsets))
    hit = cache.get(key)
    if hit is not None:
        return hit[0]
    dev = target.device
    ctx = context(dev)
    tco = cell_offsets(target)
    n = target.interval_count
    cells = int(tco[-1].item()) if n else 0
    tile = int(_lib.lib().stx_transfer_tile_cells())
    ntiles = (cells + tile - 1) // tile
    ord_ = target.interval_row_ordinals() if n else _core.empty_i32(1, dev)
    counts = _core.empty_i32(max(n, 1), dev)
    seg_off = _core.empty_i32(n + 1, dev)
    sp, sbytes, ep = ctx.scratch(_lib.lib().stx_scan_scratch_bytes(n))
    tg = _geom(target, tco)
    dg = _geom(dst.interval_set, dst.interval_cell_offsets)
    sgm = _geom(src.interval_set, src.interval_cell_offsets)
    check(_lib.lib().stx_transfer_seg_count(mode, *tg, ptr(ord_), n, *dg, *sgm, ptr(counts), ptr(seg_off),
                                            ctx.slot(0), ctx.status_ptr, sp, sbytes, ep, ctx.stream),
          "transfer_seg_count")
    (nseg,) = _core.read_counts(ctx, ctx.words[1])
    segs = _core.empty_i32(8 * max(nseg, 1), dev)
    plan = _core.empty_i32(2 * max(ntiles, 1), dev)
    check(_lib.lib().stx_transfer_seg_build(mode, *tg, ptr(ord_), n, *dg, *sgm, ptr(seg_off), nseg, ptr(segs), cells,
                                            ptr(plan), ntiles, ctx.stream), "transfer_seg_build")
    out = (plan, ntiles, cells, segs)
    cache[key] = (out, dst.interval_set, src.interval_set, dst.interval_cell_offsets, src.interval_cell_offsets)
    return out


_TRANSFER_SCHEDULE = "segments"


def set_transfer_schedule(name: str) -> None:
    """'segments' (segment tables, ratio 2; default) or 'rows' (lane group
    per target row, any ratio)."""
    global _TRANSFER_SCHEDULE
    if name not in ("segments", "rows"):
        raise ValueError("schedule must be 'segments' or 'rows'")
    _TRANSFER_SCHEDULE = name


def _seg_ok(dst: IntervalField, src: IntervalField, ratio: int) -> bool:
    return (_TRANSFER_SCHEDULE == "segments" and ratio == 2
            and src.values.numel() * src.values.element_size() < (1 << 32) - 256)


def prolong_into(target: IntervalSet, dst: IntervalField, src: IntervalField, ratio: int) -> None:
    """dst(y, x) = src(y // ratio, x // ratio) for every cell of ``target``
    (cells of target must lie in dst's geometry)."""
    if target.interval_count == 0:
        return
    if dst.values.dtype != src.values.dtype:
        raise TypeError("source and destination fields must share a dtype")
    ctx = context(dst.values.device)
    if _seg_ok(dst, src, ratio):
        plan, ntiles, cells, segs = transfer_plan(0, target, dst, src)
        check(_lib.lib().stx_transfer_seg(0, _dtype_code(dst.values), 0, ptr(plan), ntiles, cells, ptr(segs),
                                          ptr(src.values), int(src.values.numel()), ptr(dst.values), ctx.stream),
              "prolong_into")
        return
    tco = cell_offsets(target)
    n = int(dst.values.numel())
    check(_lib.lib().stx_transfer(0, _dtype_code(dst.values), 0, _group(n, target.row_count), int(ratio),
                                  *_geom(target, tco), *_geom(dst.interval_set, dst.interval_cell_offsets),
                                  *_geom(src.interval_set, src.interval_cell_offsets), ptr(src.values),
                                  ptr(dst.values), ctx.stream), "prolong_into")


def restrict_into(target: IntervalSet, dst: IntervalField, src: IntervalField, ratio: int, *,
                  accumulate: str = "f64") -> None:
    """dst(y, x) = mean of src(y*R + dy, x*R + dx) over dy, dx < R (children
    in dy-major order) for every cell of ``target``."""
    if target.interval_count == 0:
        return
    if dst.values.dtype != src.values.dtype:
        raise TypeError("source and destination fields must share a dtype")
    ctx = context(dst.values.device)
    acc64 = 1 if accumulate == "f64" else 0
    if _seg_ok(dst, src, ratio):
        plan, ntiles, cells, segs = transfer_plan(1, target, dst, src)
        check(_lib.lib().stx_transfer_seg(1, _dtype_code(dst.values), acc64, ptr(plan), ntiles, cells, ptr(segs),
                                          ptr(src.values), int(src.values.numel()), ptr(dst.values), ctx.stream),
              "restrict_into")
        return
    tco = cell_offsets(target)
    check(_lib.lib().stx_transfer(1, _dtype_code(dst.values), acc64, _group(int(src.values.numel()), target.row_count),
                                  int(ratio), *_geom(target, tco), *_geom(dst.interval_set, dst.interval_cell_offsets),
                                  *_geom(src.interval_set, src.interval_cell_offsets), ptr(src.values),
                                  ptr(dst.values), ctx.stream), "restrict_into")


def gradient_tag(field: IntervalField, threshold: float) -> IntervalSet:
    """Runs of cells with |grad u| >= threshold (central differences where
    both neighbours are in the set, one-sided otherwise), rows of the field."""
    s = field.interval_set
    dev = field.values.device
    ctx = context(dev)
    n = int(field.values.numel())
    cap = max(n // 2 + s.interval_count + 1, 1)
    out_off = _core.empty_i32(s.row_count + 1, dev)
    ob = _core.empty_i32(cap, dev)
    oe = _core.empty_i32(cap, dev)
    g = _group(n, s.row_count)
    sp, sbytes, ep = ctx.scratch(_lib.lib().stx_setop_scratch_bytes(s.row_count, 4))
    check(_lib.lib().stx_grad_tag(_dtype_code(field.values), g, ptr(s.rows), ptr(s.row_offsets), s.row_count,
                                  ptr(s.begin), ptr(s.end), ptr(field.interval_cell_offsets), ptr(field.values), float(threshold), ptr(out_off), ptr(ob), ptr(oe), cap, ctx.slot(0),
                                  ctx.status_ptr, sp, sbytes, ep, ctx.stream), "grad_tag")
    (total,) = _core.read_counts(ctx, ctx.words[1])
    return IntervalSet._trusted(ob[:total], oe[:total], out_off, s.rows, dense=s.is_dense_rows())


def upwind_batched(field: IntervalField, values: torch.Tensor, out: torch.Tensor, row_inst: torch.Tensor,
                   inst_w: torch.Tensor, inst_coef: torch.Tensor, plan=None) -> torch.Tensor:
    """One upwind step over a stack of instances (see stx_upwind_batched);
    ``plan`` from interval_stencil.build_stencil_plan(set, co, row_inst,
    inst_w) selects the segment kernel."""
    from .interval_stencil import _segment_kernel_ok, stencil_schedule

    s = field.interval_set
    ctx = context(values.device)
    if plan is not None and stencil_schedule() == "tiled" and _segment_kernel_ok(values, out):
        p, ntiles, cells, segs = plan
        check(_lib.lib().stx_upwind_seg(_dtype_code(values), ptr(p), ntiles, cells, ptr(segs), ptr(values), ptr(out),
                                        0.0, 0.0, 0.0, 1.0, 1.0, ptr(inst_coef), ctx.stream), "upwind_batched")
        return out
    check(_lib.lib().stx_upwind_batched(_dtype_code(values), _group(int(values.numel()), s.row_count), ptr(s.rows),
                                        ptr(s.row_offsets), s.row_count, ptr(s.begin), ptr(s.end),
                                        ptr(field.interval_cell_offsets), ptr(values), ptr(out), ptr(row_inst),
                                        ptr(inst_w), ptr(inst_coef), ctx.stream), "upwind_batched")
    return out
