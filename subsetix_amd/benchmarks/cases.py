"""The benchmark cases (names, geometries, operations and input-interval
counts of subsetix_cupy/benchmarks/cases.py:549-846).  Inputs are built
when a case is set up, not at import.

Geometries: SMALL 256 x 256 and LARGE 8192 x 8192 grids of 3 x 3 squares
(80 % fill) for set algebra, morphology, interval fields and multilevel
transfers (ratio 2); 512^2 / 2048^2 single squares (65 % fill) for the
stencil and VTK export.
"""

from __future__ import annotations

import os
import tempfile
from typing import Callable, Dict, List, Tuple

from .. import (
    MultiLevel2D,
    MultiLevelField2D,
    create_interval_field,
    evaluate,
    locate_interval_cells,
    make_complement,
    make_difference,
    make_input,
    make_intersection,
    make_symmetric_difference,
    make_union,
    prolong_field,
    prolong_level_field,
    prolong_level_sets,
    prolong_set,
    restrict_field,
    restrict_level_field,
    restrict_level_sets,
    restrict_set,
    step_upwind_interval,
)
from ..expressions import CuPyWorkspace, _align_interval_sets
from ..export_vtk import write_unstructured_quads_vtu
from ..morphology import clip_interval_set, dilate_interval_set, erode_interval_set, full_interval_set, ghost_zones
from ..multilevel import coarse_only, covered_by_fine
from . import BenchmarkCase, BenchmarkTarget
from .geometry import (
    GeometrySpec,
    deterministic_interval_field,
    random_interval_set,
    square_grid_interval_set,
    staircase_interval_set,
)

SMALL_SPEC = GeometrySpec(rows=256, width=256, tiles_x=3, tiles_y=3, fill_ratio=0.8)
LARGE_SPEC = GeometrySpec(rows=8192, width=8192, tiles_x=3, tiles_y=3, fill_ratio=0.8)
STENCIL_SMALL = GeometrySpec(rows=512, width=512, tiles_x=1, tiles_y=1, fill_ratio=0.65)
STENCIL_LARGE = GeometrySpec(rows=2048, width=2048, tiles_x=1, tiles_y=1, fill_ratio=0.65)
RATIO = 2


def _n(s) -> int:
    return int(s.interval_count)


def _case(name: str, description: str, build: Callable[[], Tuple[Callable, int, int, Dict]]) -> BenchmarkCase:
    def setup(_torch) -> BenchmarkTarget:
        func, repeat, warmup, meta = build()
        return BenchmarkTarget(func=func, repeat=repeat, warmup=warmup, metadata=meta)

    return BenchmarkCase(name=name, description=description, setup=setup)


def _expression(spec: GeometrySpec, op: str):
    def build():
        sx1 = int(round(spec.width * (1.0 / max(1, spec.tiles_x * 2))))
        sx2 = int(round(spec.width * (0.5 / max(1, spec.tiles_x))))
        sy2 = int(round(spec.rows * (0.5 / max(1, spec.tiles_y))))
        lhs = square_grid_interval_set(spec=spec)
        rhs = square_grid_interval_set(spec=spec, shift_x=sx1)
        a, b = make_input(lhs), make_input(rhs)
        n_in = _n(lhs) + _n(rhs)
        if op == "union":
            expr = make_union(a, b)
        elif op == "difference":
            expr = make_difference(a, b)
        elif op == "symmetric_difference":
            expr = make_symmetric_difference(a, b)
        elif op == "complement":
            uni = full_interval_set(spec.width, spec.rows)
            expr = make_complement(make_input(uni), a)
            n_in = _n(lhs) + _n(uni)
        elif op == "chain":
            extra = square_grid_interval_set(spec=spec, shift_x=sx2, shift_y=sy2)
            c = make_input(extra)
            expr = make_union(make_difference(make_intersection(a, b), c), make_intersection(b, c))
            n_in += _n(extra)
        else:
            raise ValueError(op)
        ws = CuPyWorkspace()
        meta = {"rows": spec.rows, "width": spec.width, "tiles_x": spec.tiles_x, "tiles_y": spec.tiles_y,
                "operation": op, "input_intervals": n_in}
        return (lambda: evaluate(expr, ws)), 100, 10, meta

    return build


def _align(spec: GeometrySpec):
    def build():
        lhs = random_interval_set(rows=spec.rows, width=spec.width, intervals_per_row=(1, max(2, spec.tiles_x * 2)),
                                  seed=123)
        rhs = random_interval_set(rows=spec.rows, width=spec.width, intervals_per_row=(1, max(2, spec.tiles_x * 3)),
                                  seed=987)
        meta = {"rows": spec.rows, "width": spec.width, "lhs_rows": lhs.row_count, "rhs_rows": rhs.row_count,
                "operation": "align_interval_sets", "input_intervals": _n(lhs) + _n(rhs)}
        return (lambda: _align_interval_sets(lhs, rhs)), 120, 10, meta

    return build


def _morph(spec: GeometrySpec, hx: int, hy: int, op: str):
    def build():
        g = square_grid_interval_set(spec=spec)
        w, h = spec.width, spec.rows
        funcs = {
            "dilate": lambda: clip_interval_set(dilate_interval_set(g, halo_x=hx, halo_y=hy), width=w, height=h),
            "ghost": lambda: ghost_zones(g, halo_x=hx, halo_y=hy, width=w, height=h),
            "erode": lambda: erode_interval_set(g, halo_x=hx, halo_y=hy, width=w, height=h),
        }
        meta = {"width": w, "rows": h, "halo_x": hx, "halo_y": hy, "intervals": _n(g), "operation": op,
                "input_intervals": _n(g), "boundary": "zero_exterior"}
        return funcs[op], 50, 5, meta

    return build


def _interval_field(spec: GeometrySpec, op: str):
    import torch

    def build():
        g = square_grid_interval_set(spec=spec)
        if op == "create":
            meta = {"rows": spec.rows, "width": spec.width, "operation": "interval_field_create",
                    "input_intervals": _n(g)}
            return (lambda: create_interval_field(g, fill_value=0.0, dtype=torch.float32)), 60, 5, meta
        f = deterministic_interval_field(g, width=spec.width)
        # the first (up to) 512 active cells in storage order
        from ..export_vtk import interval_cells

        r, c, _ = interval_cells(f)
        k = min(512, r.size)
        qr = torch.from_numpy(r[:k].copy()).to(g.device)
        qc = torch.from_numpy(c[:k].copy()).to(g.device)
        out = torch.empty(k, dtype=torch.int32, device=g.device)
        meta = {"rows": spec.rows, "width": spec.width, "operation": "interval_field_locate", "input_intervals": k,
                "samples": k}
        return (lambda: locate_interval_cells(f, qr, qc, out=out)), 120, 15, meta

    return build


def _multilevel(spec: GeometrySpec, op: str):
    def build():
        base = square_grid_interval_set(spec=spec)
        fine = prolong_set(base, RATIO)
        fb = deterministic_interval_field(base, width=spec.width)
        ff = prolong_field(fb, RATIO)
        nb, nf = _n(base), _n(fine)
        cb, cf = int(fb.values.numel()), int(ff.values.numel())
        ml_base = MultiLevel2D.create(2, base_ratio=RATIO)
        ml_base.set_level(0, base)
        ml_full = MultiLevel2D.create(2, base_ratio=RATIO)
        ml_full.set_level(0, base)
        ml_full.set_level(1, fine)
        mf_base = MultiLevelField2D.empty_like(ml_base)
        mf_base.set_level_field(0, fb)
        mf_full = MultiLevelField2D.empty_like(ml_full)
        mf_full.set_level_field(0, fb)
        mf_full.set_level_field(1, ff)

        def two_levels():
            ml = MultiLevel2D.create(2, base_ratio=RATIO)
            ml.set_level(0, base)
            ml.set_level(1, fine)
            return ml

        def prolong_sets():
            ml_base.set_level(0, base)
            ml_base.set_level(1, None)
            prolong_level_sets(ml_base, 0)

        def restrict_sets():
            ml_full.set_level(0, base)
            ml_full.set_level(1, fine)
            restrict_level_sets(ml_full, 0)

        def prolong_fields():
            ml_base.set_level(0, base)
            mf_base.set_level_field(0, fb)
            mf_base.set_level_field(1, None)
            prolong_level_field(ml_base, mf_base, 0)

        def restrict_fields():
            ml_full.set_level(0, base)
            ml_full.set_level(1, fine)
            mf_full.set_level_field(0, fb)
            mf_full.set_level_field(1, ff)
            restrict_level_field(ml_full, mf_full, 0)

        ops = {
            "prolong_set": (lambda: prolong_set(base, RATIO), "IntervalSet", nb),
            "restrict_set": (lambda: restrict_set(fine, RATIO), "IntervalSet", nf),
            "prolong_field": (lambda: prolong_field(fb, RATIO), "IntervalField", cb),
            "restrict_field": (lambda: restrict_field(ff, RATIO), "IntervalField", cf),
            "covered_by_fine": (lambda: covered_by_fine(two_levels(), 0), "IntervalSet", nf),
            "coarse_only": (lambda: coarse_only(base, two_levels(), 0), "IntervalSet", nb + nf),
            "prolong_level_sets": (prolong_sets, "MultiLevel2D", nb),
            "restrict_level_sets": (restrict_sets, "MultiLevel2D", nf),
            "prolong_level_field": (prolong_fields, "MultiLevelField2D", cb),
            "restrict_level_field": (restrict_fields, "MultiLevelField2D", cf),
        }
        func, entity, n_in = ops[op]
        meta = {"entity": entity, "input_intervals": n_in, "base_rows": base.row_count, "base_intervals": nb,
                "ratio": RATIO}
        return func, 40, 5, meta

    return build


def _vtu(name: str, spec: GeometrySpec):
    import torch

    def build():
        bs = square_grid_interval_set(spec=spec)
        bf = deterministic_interval_field(bs, width=spec.width)
        fs = prolong_set(bs, RATIO)
        ffield = prolong_field(bf, RATIO)
        gs = ghost_zones(fs, halo_x=1, halo_y=1, width=spec.width * RATIO, height=spec.rows * RATIO)
        gf = create_interval_field(gs, fill_value=-1.0, dtype=torch.float32)
        dx0, dy0 = 1.0 / max(1, spec.width), 1.0 / max(1, spec.rows)
        cells = [(bf, 0, dx0, dy0, 0.0, 0.0, 0), (ffield, 1, dx0 / RATIO, dy0 / RATIO, 0.0, 0.0, 0),
                 (gf, 1, dx0 / RATIO, dy0 / RATIO, 0.0, 0.0, 1)]
        path = os.path.join(tempfile.gettempdir(), "subsetix_bench_vtu", f"{name}.vtu")
        n_in = int(bf.values.numel() + ffield.values.numel() + gf.values.numel())
        meta = {"rows": spec.rows, "width": spec.width, "levels": 2, "operation": "write_unstructured_quads_vtu",
                "input_intervals": n_in}
        return (lambda: write_unstructured_quads_vtu(path, cells)), 12, 3, meta

    return build


def _stencil(spec: GeometrySpec, variant: str):
    import torch

    def build():
        square = square_grid_interval_set(spec=spec)
        s = square if variant == "square" else staircase_interval_set(spec=spec, reference=square)
        if variant != "square" and not torch.equal(s.row_offsets, square.row_offsets):
            raise RuntimeError("staircase changed the row structure")
        f = deterministic_interval_field(s, width=spec.width)
        out = torch.empty_like(f.values)
        kw = dict(width=spec.width, height=spec.rows, a=0.35, b=-0.25, dt=0.01, dx=1.0 / max(1, spec.width),
                  dy=1.0 / max(1, spec.rows), out=out)
        for _ in range(3):
            step_upwind_interval(f, **kw)
        n = int(f.values.numel())
        meta = {"entity": "IntervalField", "rows": spec.rows, "width": spec.width, "active_cells": n, "a": 0.35,
                "b": -0.25, "dt": 0.01, "dx": kw["dx"], "dy": kw["dy"], "boundary": "zero_exterior",
                "variant": variant, "input_intervals": n}
        return (lambda: step_upwind_interval(f, **kw)), 40, 5, meta

    return build


def _sized(kind: str, desc: str, make, specs=(("small", SMALL_SPEC), ("large", LARGE_SPEC))) -> List[BenchmarkCase]:
    return [_case(f"{kind}_{size}", f"{desc} ({size} domain)", make(spec)) for size, spec in specs]


_STENCIL_SPECS = (("small", STENCIL_SMALL), ("large", STENCIL_LARGE))

CASES: List[BenchmarkCase] = (
    _sized("expr_union", "Interval union, 3x3 squares", lambda s: _expression(s, "union"))
    + _sized("expr_chain", "(A n B) - C u (B n C), 3x3 squares", lambda s: _expression(s, "chain"))
    + _sized("expr_difference", "Interval difference, 3x3 squares", lambda s: _expression(s, "difference"))
    + _sized("expr_symmetric_difference", "Symmetric difference, 3x3 squares",
             lambda s: _expression(s, "symmetric_difference"))
    + _sized("expr_complement", "Complement in the full domain", lambda s: _expression(s, "complement"))
    + _sized("expr_align", "Row alignment of two random sets", _align)
    + _sized("morph_dilate", "Dilate (4, 2) + clip", lambda s: _morph(s, 4, 2, "dilate"))
    + _sized("morph_ghost", "Ghost zones (3, 3)", lambda s: _morph(s, 3, 3, "ghost"))
    + _sized("morph_erode", "Erode (2, 2)", lambda s: _morph(s, 2, 2, "erode"))
    + _sized("interval_field_create", "IntervalField allocation", lambda s: _interval_field(s, "create"))
    + _sized("interval_field_locate", "Point location of 512 cells", lambda s: _interval_field(s, "locate"))
    + [c for op in ("prolong_set", "restrict_set", "prolong_field", "restrict_field", "covered_by_fine",
                    "coarse_only", "prolong_level_sets", "restrict_level_sets", "prolong_level_field",
                    "restrict_level_field")
       for c in _sized(f"multilevel_{op}", f"Multilevel {op.replace('_', ' ')}, ratio 2",
                       lambda s, op=op: _multilevel(s, op))]
    + [_case(f"export_vtu_{size}", f"VTU export of two levels + ghosts ({size} domain)", _vtu(f"export_vtu_{size}", s))
       for size, s in _STENCIL_SPECS]
    + _sized("stencil_interval_square", "Upwind step on a square", lambda s: _stencil(s, "square"), _STENCIL_SPECS)
    + _sized("stencil_interval_stair", "Upwind step on a staircase", lambda s: _stencil(s, "staircase"),
             _STENCIL_SPECS)
)

__all__ = ["CASES", "SMALL_SPEC", "LARGE_SPEC", "STENCIL_SMALL", "STENCIL_LARGE"]
