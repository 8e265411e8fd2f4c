"""Level-to-level kernels on interval fields: fused transfers onto a target
subset, gradient tagging on sparse levels and the batched upwind step.

These are the building blocks of the sparse multi-level drivers
(``subsetix_amd.amr2.composite`` and ``subsetix_amd.batch``): unlike the
reference's synchronisation (restrict the whole fine field, gather, scatter;
fields.py:570-612) each transfer is one pass over the target cells.
"""

from __future__ import annotations

import torch

from . import _core, _lib, autotune
from ._lib import _F64, _I32, _I64, _P, check, context, ptr
from .expressions import IntervalSet
from .interval_field import IntervalField, cell_offsets, known_cell_count, remember_cell_count

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
_lib.register("stx_grad_tag_seg_words", _I64, [_I64])
_lib.register("stx_grad_tag_seg", _I32, [_I32, _P, _I64, _I64, _P, _P, _P, _I64, _P, _F64, _P, _I64, _P, _P, _P, _I64,
                                         _P, _P, _P, _I64, _I32, _P])


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
    """Segment table of a transfer onto ``target`` (mode 0: ratio-2 prolong,
    1: ratio-2 restrict, 2: ratio-1 copy), cached on the target set, keyed by
    the destination and source geometries."""
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
    known = known_cell_count(target, tco)
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
    if known is None:  # the cell count rides on the segment count's read
        nseg, cells = _core.read_counts(ctx, ctx.words[1], tco[-1])
        remember_cell_count(target, tco, cells)
    else:
        (nseg,), cells = _core.read_counts(ctx, ctx.words[1]), known
    tile = int(_lib.lib().stx_transfer_tile_cells())
    ntiles = (cells + tile - 1) // tile
    segs = _core.empty_i32(8 * max(nseg, 1), dev)
    plan = _core.empty_i32(2 * max(ntiles, 1), dev)
    check(_lib.lib().stx_transfer_seg_build(mode, *tg, ptr(ord_), n, *dg, *sgm, ptr(seg_off), nseg, ptr(segs), cells,
                                            ptr(plan), ntiles, ctx.stream), "transfer_seg_build")
    out = (plan, ntiles, cells, segs)
    cache[key] = (out, dst.interval_set, src.interval_set, dst.interval_cell_offsets, src.interval_cell_offsets)
    return out


_TRANSFER_SCHEDULE = "auto"


def set_transfer_schedule(name: str) -> None:
    """'auto' (default: the autotuner's choice when it has one, else segment
    tables where they apply), 'segments' (segment tables: ratio 2, and the
    ratio-1 prolong) or 'rows' (lane group per target row, any ratio)."""
    global _TRANSFER_SCHEDULE
    if name not in ("auto", "segments", "rows"):
        raise ValueError("schedule must be 'auto', 'segments' or 'rows'")
    _TRANSFER_SCHEDULE = name


def _seg_mode(mode: int, ratio: int):
    """Segment-table mode of a transfer, None where tables do not apply: the
    ratio-2 prolong (0) and restrict (1), and the ratio-1 prolong as a copy
    (2).  A ratio-1 restrict reads absent sources as 0 and stays on rows."""
    if ratio == 2:
        return mode
    if ratio == 1 and mode == 0:
        return 2
    return None


def _seg_possible(src: IntervalField, mode: int, ratio: int) -> bool:
    return (_seg_mode(mode, ratio) is not None
            and src.values.numel() * src.values.element_size() < (1 << 32) - 256)


def _transfer(mode: int, target: IntervalSet, dst: IntervalField, src: IntervalField, ratio: int, acc64: int,
              schedule=None) -> None:
    """One fused transfer with the schedule chosen as for the stencil:
    ``schedule`` ("rows" for a one-shot transfer, where building a segment
    table costs more than it saves), else the module's forced schedule, else
    the autotuner's (candidates run on a copy of the destination), else
    segment tables when they apply."""
    ctx = context(dst.values.device)
    dcode = _dtype_code(dst.values)
    seg_ok = _seg_possible(src, mode, ratio)
    smode = _seg_mode(mode, ratio)
    group_cells = int(dst.values.numel()) if mode == 0 else int(src.values.numel())

    def launch(choice, d: IntervalField) -> None:
        if choice == "seg":
            plan, ntiles, cells, segs = transfer_plan(smode, target, d, src)
            check(_lib.lib().stx_transfer_seg(smode, dcode, acc64, ptr(plan), ntiles, cells, ptr(segs),
                                              ptr(src.values), int(src.values.numel()), ptr(d.values), ctx.stream),
                  "prolong_into" if mode == 0 else "restrict_into")
            return
        tco = cell_offsets(target)
        check(_lib.lib().stx_transfer(mode, dcode, acc64, int(choice), int(ratio), *_geom(target, tco),
                                      *_geom(d.interval_set, d.interval_cell_offsets),
                                      *_geom(src.interval_set, src.interval_cell_offsets), ptr(src.values),
                                      ptr(d.values), ctx.stream), "prolong_into" if mode == 0 else "restrict_into")

    fixed = "seg" if (seg_ok and _TRANSFER_SCHEDULE != "rows" and schedule != "rows") else _group(
        group_cells, target.row_count)
    choice = fixed
    if _TRANSFER_SCHEDULE == "auto" and schedule is None:
        def factory():
            scratch = IntervalField(dst.interval_set, dst.values.clone(), dst.interval_cell_offsets)
            cands = (("seg",) if seg_ok else ()) + autotune.GROUPS
            return {c: (lambda c=c: launch(c, scratch)) for c in cands}

        key = ("transfer", mode, dcode, acc64, int(ratio), seg_ok, autotune.bucket(target.interval_count),
               autotune.bucket(group_cells))
        tuned = autotune.pick(key, factory)
        choice = fixed if tuned is None else tuned
    launch(choice, dst)


def prolong_into(target: IntervalSet, dst: IntervalField, src: IntervalField, ratio: int, *,
                 schedule=None) -> None:
    """dst(y, x) = src(y // ratio, x // ratio) for every cell of ``target``
    (cells of target must lie in dst's geometry).  ``schedule="rows"``: no
    segment table (a transfer made once)."""
    if target.interval_count == 0:
        return
    if dst.values.dtype != src.values.dtype:
        raise TypeError("source and destination fields must share a dtype")
    _transfer(0, target, dst, src, int(ratio), 0, schedule)


def restrict_into(target: IntervalSet, dst: IntervalField, src: IntervalField, ratio: int, *,
                  accumulate: str = "f64") -> None:
    """dst(y, x) = mean of src(y*R + dy, x*R + dx) over dy, dx < R (children
    in dy-major order) for every cell of ``target``."""
    if target.interval_count == 0:
        return
    if dst.values.dtype != src.values.dtype:
        raise TypeError("source and destination fields must share a dtype")
    _transfer(1, target, dst, src, int(ratio), 1 if accumulate == "f64" else 0)


def gradient_tag(field: IntervalField, threshold: float, plan=None) -> IntervalSet:
    """Runs of cells with |grad u| >= threshold (central differences where
    both neighbours are in the set, one-sided otherwise), rows of the field.
    ``plan``: the field's stencil segment table (interval_stencil.
    build_stencil_plan) for widths that contain every cell of the set; it
    enables the segment schedule (bit masks, no row walks)."""
    from .interval_stencil import _segment_kernel_ok

    s = field.interval_set
    dev = field.values.device
    ctx = context(dev)
    n = int(field.values.numel())
    cap = max(n // 2 + s.interval_count + 1, 1)
    out_off = _core.empty_i32(s.row_count + 1, dev)
    ob = _core.empty_i32(cap, dev)
    oe = _core.empty_i32(cap, dev)
    dcode = _dtype_code(field.values)
    seg_ok = plan is not None and plan[2] == n and _segment_kernel_ok(field.values, field.values)

    def launch(g) -> None:
        if g == "seg":
            p, ntiles, ncells, segs = plan
            nw = int(_lib.lib().stx_grad_tag_seg_words(ncells))
            words = _core.empty_i32(6 * nw + 1, dev)  # tag, lex, start, end, count, prefix (+ total)
            sp, sbytes, ep = ctx.scratch(_lib.lib().stx_scan_scratch_bytes(nw))
            check(_lib.lib().stx_grad_tag_seg(dcode, ptr(p), ntiles, ncells, ptr(segs), ptr(s.row_offsets),
                                              ptr(field.interval_cell_offsets), s.row_count, ptr(field.values),
                                              float(threshold), ptr(words), nw, ptr(out_off), ptr(ob), ptr(oe), cap,
                                              ctx.slot(0), ctx.status_ptr, sp, sbytes, ep, ctx.stream), "grad_tag")
            return
        sp, sbytes, ep = ctx.scratch(_lib.lib().stx_setop_scratch_bytes(s.row_count, g))
        check(_lib.lib().stx_grad_tag(dcode, int(g), ptr(s.rows), ptr(s.row_offsets), s.row_count, ptr(s.begin),
                                      ptr(s.end), ptr(field.interval_cell_offsets), ptr(field.values),
                                      float(threshold), ptr(out_off), ptr(ob), ptr(oe), cap, ctx.slot(0),
                                      ctx.status_ptr, sp, sbytes, ep, ctx.stream), "grad_tag")

    cands = (("seg",) if seg_ok else ()) + autotune.GROUPS
    key = ("grad_tag", dcode, seg_ok, autotune.bucket(n), autotune.bucket(n / max(s.row_count, 1)))
    g = autotune.pick(key, lambda: {c: (lambda c=c: launch(c)) for c in cands})
    launch(("seg" if seg_ok else _group(n, s.row_count)) if g is None else g)
    (total,) = _core.read_counts(ctx, ctx.words[1])
    return IntervalSet._trusted(ob[:total], oe[:total], out_off, s.rows, dense=s.is_dense_rows())


def upwind_batched(field: IntervalField, values: torch.Tensor, out: torch.Tensor, row_inst: torch.Tensor,
                   inst_w: torch.Tensor, inst_coef: torch.Tensor, plan=None) -> torch.Tensor:
    """One upwind step over a stack of instances (see stx_upwind_batched);
    ``plan`` from interval_stencil.build_stencil_plan(set, co, row_inst,
    inst_w) selects the segment kernel."""
    from . import interval_stencil
    from .interval_stencil import _segment_kernel_ok

    sched = interval_stencil.stencil_schedule()
    s = field.interval_set
    ctx = context(values.device)
    dcode = _dtype_code(values)
    cells = int(values.numel())

    def launch(choice) -> None:
        if choice == "seg":
            p, ntiles, ncells, segs = plan
            check(_lib.lib().stx_upwind_seg(dcode, ptr(p), ntiles, ncells, ptr(segs), ptr(values), ptr(out), 0.0, 0.0,
                                            0.0, 1.0, 1.0, ptr(inst_coef), ctx.stream), "upwind_batched")
        else:
            check(_lib.lib().stx_upwind_batched(dcode, int(choice), ptr(s.rows), ptr(s.row_offsets), s.row_count,
                                                ptr(s.begin), ptr(s.end), ptr(field.interval_cell_offsets),
                                                ptr(values), ptr(out), ptr(row_inst), ptr(inst_w), ptr(inst_coef),
                                                ctx.stream), "upwind_batched")

    seg_ok = plan is not None and _segment_kernel_ok(values, out)
    fixed = "seg" if (seg_ok and sched != "rows") else _group(cells, s.row_count)
    choice = fixed
    if sched == "auto":
        cands = (("seg",) if seg_ok else ()) + autotune.GROUPS
        key = ("upwind", dcode, True, seg_ok, autotune.bucket(cells), autotune.bucket(cells / max(s.row_count, 1)))
        tuned = autotune.pick(key, lambda: {c: (lambda c=c: launch(c)) for c in cands})
        choice = fixed if tuned is None else tuned
    launch(choice)
    return out
