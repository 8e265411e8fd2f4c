"""Import name ``subsetix_cupy`` for existing callers of the reference
(subsetix_cupy/__init__.py:3-87): the package and its modules
(``subsetix_cupy.expressions``, ``.interval_field``, ``.multilevel``,
``.morphology``, ``.interval_stencil``, ``.export_vtk``, ``.benchmarks``) are
the modules of :mod:`subsetix_amd`, so ``from subsetix_cupy.expressions
import IntervalSet`` keeps working unchanged on the gfx950 kernels."""

import importlib
import sys

import subsetix_amd as _impl
from subsetix_amd import *  # noqa: F401,F403
from subsetix_amd import __all__  # noqa: F401

for _name in ("expressions", "interval_field", "multilevel", "morphology", "interval_stencil", "export_vtk",
              "benchmarks"):
    sys.modules[f"{__name__}.{_name}"] = importlib.import_module(f"{_impl.__name__}.{_name}")
del _name
