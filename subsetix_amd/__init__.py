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
]
