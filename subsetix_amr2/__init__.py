"""Import name ``subsetix_amr2`` for existing callers of the reference
(subsetix_amr2/__init__.py:9-62): the package and its modules
(``subsetix_amr2.api``, ``.runner``, ``.fields``, ``.geometry``,
``.regrid``, ``.simulation``, ``.export``) are the modules of
:mod:`subsetix_amd.amr2`."""

import importlib
import sys

import subsetix_amd.amr2 as _impl
from subsetix_amd.amr2 import *  # noqa: F401,F403
from subsetix_amd.amr2 import __all__  # noqa: F401

for _name in ("api", "runner", "fields", "geometry", "regrid", "simulation", "export", "mass"):
    sys.modules[f"{__name__}.{_name}"] = importlib.import_module(f"{_impl.__name__}.{_name}")
del _name
