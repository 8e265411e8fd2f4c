"""Thin host wrappers over the C-ABI (tensors in, tensors out).

Every function here launches kernels of libsubsetix_hip.so on the current
stream of the tensors' device and, where the reference API needs an exact
output size, performs ONE device->host read of the produced counts.
"""

from __future__ import annotations

import ctypes
from typing import Optional, Tuple

import torch

from . import _lib, autotune
from ._lib import _I32, _I64, _P, check, context, ptr

I32 = torch.int32

_lib.register("stx_validate_csr", _I32, [_P, _I64, _P, _I64, _P, _P])
_lib.register("stx_interval_rows", _I32, [_P, _I64, _P, _I64, _P, _P, _P])
_lib.register("stx_rows_lookup", _I32, [_P, _I64, _I32, _I64, _P, _P])
_lib.register("stx_dense_offsets", _I32, [_P, _P, _I64, _I64, _I64, _P, _P])
_lib.register("stx_compact_rows", _I32, [_P, _P, _I64, _I32, _P, _P, _P, _P, _P, _P, _I64, _I32, _P])
_lib.register("stx_coarse_rows", _I32, [_P, _P, _I64, _I32, _I32, _P, _P, _I64, _P, _P, _P, _P, _I64, _I32, _P])
_lib.register("stx_prolong_set", _I32, [_P, _P, _I64, _P, _P, _I64, _I32, _P, _P, _P, _P, _P, _P])
_lib.register("stx_prolong_rows", _I32, [_P, _I64, _I32, _P, _P])
_lib.register("stx_fill_i32", _I32, [_P, _I64, _I32, _I32, _P])
_lib.register("stx_add_i32", _I32, [_P, _P, _I64, _I32, _P])

XF_NONE, XF_WIDEN, XF_COARSEN, XF_CLIP, XF_SCALE = 0, 1, 2, 3, 4
OP_CODES = {"union": 0, "intersection": 1, "difference": 2}


def empty_i32(n: int, device) -> torch.Tensor:
    return torch.empty(max(int(n), 0), dtype=I32, device=device)


def zeros_i32(n: int, device) -> torch.Tensor:
    return torch.zeros(max(int(n), 0), dtype=I32, device=device)


def fill_i32(t: torch.Tensor, value: int, step: int = 0) -> torch.Tensor:
    ctx = context(t.device)
    check(_lib.lib().stx_fill_i32(ptr(t), t.numel(), int(value), int(step), ctx.stream), "fill_i32")
    return t


def arange_i32(n: int, device, start: int = 0) -> torch.Tensor:
    return fill_i32(empty_i32(n, device), start, 1)


def full_i32(n: int, value: int, device) -> torch.Tensor:
    return fill_i32(empty_i32(n, device), value, 0)


def add_i32(src: torch.Tensor, value: int) -> torch.Tensor:
    out = torch.empty_like(src)
    ctx = context(src.device)
    check(_lib.lib().stx_add_i32(ptr(src), ptr(out), src.numel(), int(value), ctx.stream), "add_i32")
    return out


def validate_csr(off: torch.Tensor, rows: torch.Tensor, n: int) -> int:
    """Returns the flag word of the CSR checks (bit0 off[0]!=0, bit1
    off[-1]!=n, bit2 rows not increasing, bit3 offsets decreasing, bit4 rows
    differ from arange)."""
    ctx = context(off.device)
    flags = torch.zeros(1, dtype=I32, device=off.device)
    check(_lib.lib().stx_validate_csr(ptr(off), rows.numel(), ptr(rows), int(n), ptr(flags), ctx.stream), "validate_csr")
    return int(flags.item())


def interval_rows(off: torch.Tensor, rows: torch.Tensor, n: int, want_row=True, want_ord=False):
    ctx = context(off.device)
    r = empty_i32(n, off.device) if want_row else None
    o = empty_i32(n, off.device) if want_ord else None
    if n:
        check(_lib.lib().stx_interval_rows(ptr(off), rows.numel(), ptr(rows), int(n), ptr(r), ptr(o), ctx.stream),
              "interval_rows")
    return r, o


def rows_lookup(rows: torch.Tensor, y0: int, h: int) -> torch.Tensor:
    ctx = context(rows.device)
    m = empty_i32(h, rows.device)
    check(_lib.lib().stx_rows_lookup(ptr(rows), rows.numel(), int(y0), int(h), ptr(m), ctx.stream), "rows_lookup")
    return m


def choose_group(n_elems: int, n_rows: int) -> int:
    """Lane-group size per output row from the mean merged row length (the
    row-bucket schedule; every size is exact for any row length, larger rows
    just take more passes)."""
    avg = n_elems / max(n_rows, 1)
    if avg >= 24:
        return 64
    if avg >= 6:
        return 16
    return 4


_GROUP_OVERRIDE: Optional[int] = None


def set_group_override(g: Optional[int]) -> None:
    global _GROUP_OVERRIDE
    _GROUP_OVERRIDE = g


_lib.register("stx_setop_noalloc", _I32, [_I32, _P, _P, _P, _I64, _P, _P, _P, _I64, _I64, _P, _P, _P, _I64, _P, _P,
                                          _P])
_lib.register("stx_setop_noalloc_max_scratch_bytes", _I64, [])


def setop_raw(op: int, a, b, *, nrows: int, d_nrows: Optional[torch.Tensor] = None,
              map_a=None, map_b=None, xf_a=(XF_NONE, 0), xf_b=(XF_NONE, 0),
              capacity: Optional[int] = None, buffers=None, group: Optional[int] = None):
    """Launch the fused set kernel.  ``a``/``b`` are (begin, end, row_offsets)
    tensor triples (``b`` may be None for op 0: single-operand merge).
    Returns (out_off[nrows+1], out_begin[cap], out_end[cap], ctx) with the
    total in ctx slot 0 (not yet read)."""
    a_beg, a_end, a_off = a
    dev = a_off.device
    ctx = context(dev)
    na = a_beg.numel()
    nb = b[0].numel() if b is not None else 0
    cap = int(capacity) if capacity is not None else max(na + nb, 1)
    out_off = empty_i32(nrows + 1, dev)
    if buffers is not None:
        ob, oe = buffers
    else:
        ob, oe = empty_i32(cap, dev), empty_i32(cap, dev)
    b_beg, b_end, b_off = b if b is not None else (None, None, None)

    def launch(g: int) -> None:
        sp, sbytes, ep = ctx.scratch(_lib.lib().stx_setop_scratch_bytes_n(nrows, na + nb, g))
        check(_lib.lib().stx_setop(int(op), ptr(a_beg), ptr(a_end), ptr(a_off), ptr(map_a), int(xf_a[0]),
                                   int(xf_a[1]), int(na), ptr(b_beg), ptr(b_end), ptr(b_off), ptr(map_b),
                                   int(xf_b[0]), int(xf_b[1]), int(nb), ptr(d_nrows), int(nrows), int(g),
                                   ptr(out_off), ptr(ob), ptr(oe), cap, ctx.slot(0), ctx.status_ptr, sp, sbytes, ep,
                                   ctx.stream), "setop")

    g = group or _GROUP_OVERRIDE or 0
    if g == 0:
        tile_ok = map_a is None and map_b is None and d_nrows is None
        plain2 = b is not None and xf_a[0] == XF_NONE and xf_b[0] == XF_NONE
        cands = (((-7, -8) if plain2 else ()) + ((-6, -5, -1, -2, -3, -4) if b is not None else (-6, -5, -1, -3, -4))
                 if tile_ok else ()) + autotune.GROUPS
        key = ("setop", int(op), tile_ok, autotune.bucket(na + nb), autotune.bucket((na + nb) / max(nrows, 1)))
        g = autotune.pick(key, lambda: {c: (lambda c=c: launch(c)) for c in cands}) or 0
        if g == 0 and not tile_ok:
            g = choose_group(na + nb, nrows)
    launch(int(g))
    return out_off, ob, oe, ctx


def rows_union(rows_a: torch.Tensor, shift_a: int, rows_b: torch.Tensor, shift_b: int):
    """Sorted union of two row-id lists (+shifts) with source maps.  Returns
    (rows_u, map_a, map_b, d_nu) sized for na+nb; the count is in d_nu."""
    dev = rows_a.device
    ctx = context(dev)
    na, nb = rows_a.numel(), rows_b.numel()
    rows_u = empty_i32(na + nb, dev)
    map_a = empty_i32(na + nb, dev)
    map_b = empty_i32(na + nb, dev)
    tmp = empty_i32(2 * (nb + 1), dev)
    d_nu = empty_i32(1, dev)
    sp, sbytes, ep = ctx.scratch(_lib.lib().stx_scan_scratch_bytes(nb))
    check(_lib.lib().stx_rows_union(ptr(rows_a), na, int(shift_a), ptr(rows_b), nb, int(shift_b), ptr(rows_u),
                                    ptr(map_a), ptr(map_b), ptr(d_nu), ptr(tmp), ctx.status_ptr, sp, sbytes, ep, ctx.stream),
          "rows_union")
    return rows_u, map_a, map_b, d_nu


def compact_rows(rows: torch.Tensor, off: torch.Tensor, shift: int = 0):
    """Drop rows without intervals.  Returns (rows', off', d_count) sized R."""
    dev = rows.device
    ctx = context(dev)
    R = rows.numel()
    rows_o = empty_i32(R, dev)
    off_o = empty_i32(R + 1, dev)
    tmp = empty_i32(2 * (R + 1), dev)
    d_cnt = empty_i32(1, dev)
    sp, sbytes, ep = ctx.scratch(_lib.lib().stx_scan_scratch_bytes(R))
    check(_lib.lib().stx_compact_rows(ptr(rows), ptr(off), R, int(shift), ptr(rows_o), ptr(off_o), ptr(d_cnt), ptr(tmp),
                                      ctx.status_ptr, sp, sbytes, ep, ctx.stream), "compact_rows")
    return rows_o, off_o, d_cnt


def coarse_rows(rows: torch.Tensor, off: torch.Tensor, ratio: int, all_rows: bool, want_maps: bool):
    dev = rows.device
    ctx = context(dev)
    R = rows.numel()
    crow = empty_i32(R, dev)
    maps = empty_i32(ratio * R, dev) if want_maps else None
    tmp = empty_i32(2 * (R + 1), dev)
    d_cnt = empty_i32(1, dev)
    sp, sbytes, ep = ctx.scratch(_lib.lib().stx_scan_scratch_bytes(R))
    check(_lib.lib().stx_coarse_rows(ptr(rows), ptr(off), R, int(ratio), int(bool(all_rows)), ptr(crow), ptr(maps), R,
                                     ptr(d_cnt), ptr(tmp), ctx.status_ptr, sp, sbytes, ep, ctx.stream), "coarse_rows")
    return crow, maps, d_cnt


def read_counts(ctx, *dev_scalars) -> list:
    """One D2H read of several device int32 scalars plus the status word."""
    if dev_scalars:
        vals = torch.cat([ctx.words[:1]] + [d.reshape(1) for d in dev_scalars]).tolist()
    else:
        vals = ctx.words[:1].tolist()
    if vals[0]:
        ctx.read_slots(0)  # raises
    return vals[1:]
