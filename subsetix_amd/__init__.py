"""subsetix_amd: MI355X-native interval-set AMR hot path.

Drop-in for the ``subsetix_cupy`` API surface (subsetix_cupy/__init__.py:3-87)
backed by hand-written gfx950 HIP kernels (libsubsetix_hip.so, ctypes C-ABI,
torch tensors / DLPack for device arrays).  ``subsetix_amd.amr2`` mirrors
``subsetix_amr2``.
"""

from .expressions import (
    CuPyWorkspace,
    Expr,
    IntervalSet,
    build_interval_set,
    evaluate,
    make_complement,
    make_difference,
    make_input,
    make_intersection,
    make_symmetric_difference,
    make_union,
)
from .interval_field import IntervalField, create_interval_field, get_cell, locate_interval_cells, set_cell
from .interval_stencil import step_upwind_interval, step_upwind_interval_field
from .morphology import (
    boundary_for_direction,
    boundary_layer,
    clip_interval_set,
    dilate_interval_set,
    erode_interval_set,
    full_interval_set,
    ghost_zones,
    interior_for_direction,
    translate_interval_set,
)
from .multilevel import (
    MultiLevel2D,
    MultiLevelField2D,
    coarse_only,
    covered_by_fine,
    prolong_field,
    prolong_level_field,
    prolong_level_sets,
    prolong_set,
    restrict_field,
    restrict_level_field,
    restrict_level_sets,
    restrict_set,
)

__all__ = [
    "Expr",
    "IntervalSet",
    "CuPyWorkspace",
    "build_interval_set",
    "evaluate",
    "make_difference",
    "make_complement",
    "make_input",
    "make_intersection",
    "make_symmetric_difference",
    "make_union",
    "IntervalField",
    "create_interval_field",
    "get_cell",
    "locate_interval_cells",
    "set_cell",
    "MultiLevel2D",
    "MultiLevelField2D",
    "prolong_set",
    "restrict_set",
    "prolong_field",
    "restrict_field",
    "prolong_level_sets",
    "restrict_level_sets",
    "prolong_level_field",
    "restrict_level_field",
    "covered_by_fine",
    "coarse_only",
    "translate_interval_set",
    "interior_for_direction",
    "boundary_for_direction",
    "boundary_layer",
    "dilate_interval_set",
    "erode_interval_set",
    "ghost_zones",
    "full_interval_set",
    "clip_interval_set",
    "step_upwind_interval",
    "step_upwind_interval_field",
]
