"""Two-level VTK snapshot (drop-in for ``subsetix_amr2.export``,
export.py:37-109): one .vtu holding, as quads,

  level 0, ghost 0: coarse-only cells away from the refine set;
  level 0, ghost 1: coarse-only cells within ``ghost_halo`` of it;
  level 1, ghost 0: the fine cells (prolongation of the refine set);
  level 1, ghost 1: the fine ring of width ``ghost_halo * ratio`` around them.

The four cell sets are computed with this package's set algebra on the
device, each field is gathered onto them by one subset-copy kernel, and the
writer (subsetix_amd.export_vtk) reads each gathered field back with one
device -> host copy.
"""

from __future__ import annotations

import os
from typing import Dict

from ..expressions import IntervalSet, binary_op
from ..export_vtk import write_unstructured_quads_vtu
from ..interval_field import IntervalField
from ..morphology import ghost_zones
from .fields import gather_interval_subset


def _require_field(f, what: str) -> IntervalField:
    if not isinstance(f, IntervalField):
        raise TypeError(f"{what}: expected an IntervalField, got {type(f).__name__}")
    return f


def save_two_level_vtk(out_dir: str, base: str, step: int, *, coarse_field, fine_field, refine_set: IntervalSet,
                       coarse_only_set: IntervalSet, fine_set: IntervalSet, dx_coarse: float, dy_coarse: float,
                       ratio: int = 2, time_value: float = 0.0, ghost_halo: int = 1, width: int | None = None,
                       height: int | None = None) -> Dict[str, str]:
    """Write ``{base}_mesh_step{step:04d}.vtu`` in ``out_dir``; returns
    ``{"mesh_vtu": file name}``.  ``width`` / ``height`` default to the coarse
    set's extent (largest interval end, row count)."""
    cf = _require_field(coarse_field, "coarse_field")
    ff = _require_field(fine_field, "fine_field")
    cs = cf.interval_set
    if height is None:
        height = cs.row_count
    if width is None:
        width = int(cs.end.max().item()) if cs.interval_count else 0
    w, h, r = int(width), int(height), int(ratio)
    halo = int(ghost_halo)
    ring = ghost_zones(refine_set, halo_x=halo, halo_y=halo, width=w, height=h)
    coarse_ghost = binary_op(ring, coarse_only_set, "intersection")
    coarse_active = binary_op(coarse_only_set, coarse_ghost, "difference")
    fine_ghost = ghost_zones(fine_set, halo_x=max(0, halo * r), halo_y=max(0, halo * r), width=w * r, height=h * r)
    dxf, dyf = dx_coarse / r, dy_coarse / r
    parts = ((cf, coarse_active, 0, dx_coarse, dy_coarse, 0), (cf, coarse_ghost, 0, dx_coarse, dy_coarse, 1),
             (ff, fine_set, 1, dxf, dyf, 0), (ff, fine_ghost, 1, dxf, dyf, 1))
    entries = []
    for field, subset, level, dx, dy, ghost in parts:
        if subset.interval_count:
            entries.append((gather_interval_subset(field, subset), level, dx, dy, 0.0, 0.0, ghost))
    os.makedirs(out_dir, exist_ok=True)
    name = f"{base}_mesh_step{int(step):04d}.vtu"
    write_unstructured_quads_vtu(os.path.join(out_dir, name), entries)
    return {"mesh_vtu": name}


__all__ = ["save_two_level_vtk"]
