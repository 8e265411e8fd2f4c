"""CPU: the C-ABI library loads and exports every function include/*.h
declares; Python bindings match the declared argument counts."""

import ctypes
import glob
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _declared():
    names = {}
    for h in glob.glob(os.path.join(ROOT, "include", "*.h")):
        text = open(h).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        for m in re.finditer(r"\b(?:int32_t|int64_t|void|const char\*)\s+(stx_\w+)\s*\(([^;]*?)\)\s*;", text, flags=re.S):
            args = m.group(2).strip()
            nargs = 0 if args in ("", "void") else args.count(",") + 1
            names[m.group(1)] = nargs
    return names


def test_header_declares_functions():
    names = _declared()
    assert "stx_setop" in names and "stx_upwind" in names
    assert len(names) >= 30


def test_library_exports_every_declared_symbol():
    from subsetix_amd import _lib

    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in _declared() if not hasattr(lib, n)]
    assert not missing, missing


def test_every_exported_symbol_is_declared():
    import subprocess

    from subsetix_amd import _lib

    out = subprocess.run(["nm", "-D", "--defined-only", _lib.LIB_PATH], capture_output=True, text=True,
                         check=True).stdout
    exported = {line.split()[-1] for line in out.splitlines() if " T stx_" in line}
    undeclared = sorted(exported - set(_declared()))
    assert not undeclared, undeclared


def test_python_bindings_match_header():
    import subsetix_amd  # noqa: F401  (registers every binding)
    import subsetix_amd.amr2  # noqa: F401
    import subsetix_amd.amr2.composite  # noqa: F401
    import subsetix_amd.batch  # noqa: F401
    import subsetix_amd.interval_stencil  # noqa: F401
    import subsetix_amd.levels  # noqa: F401
    import subsetix_amd.serve  # noqa: F401
    from subsetix_amd import _lib

    decl = _declared()
    for name, (_res, args) in _lib._SIGNATURES.items():
        assert name in decl, f"{name} bound in Python but not declared in include/"
        assert len(args) == decl[name], (name, len(args), decl[name])


def test_product_has_no_oracle_dependency():
    for path in glob.glob(os.path.join(ROOT, "subsetix_amd", "**", "*.py"), recursive=True):
        text = open(path).read()
        assert "oracle" not in re.sub(r"#.*", "", text).replace("the oracle", ""), path
