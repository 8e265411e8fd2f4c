"""CPU: the C-ABI library loads and exports every function include/*.h
declares; Python bindings match the declared argument counts."""

import ctypes
import glob
import os
import re

import pytest

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


def test_every_called_entry_point_is_registered():
    """Every lib().stx_* the package calls has a ctypes signature once the
    package is imported (an unprototyped call would pass device pointers as
    32-bit ints), and the library handle refuses unregistered names."""
    import importlib
    import pkgutil

    import subsetix_amd
    from subsetix_amd import _lib

    for m in pkgutil.walk_packages(subsetix_amd.__path__, "subsetix_amd."):
        if not m.name.endswith("libsubsetix_hip"):
            importlib.import_module(m.name)
    called = set()
    for path in glob.glob(os.path.join(ROOT, "subsetix_amd", "**", "*.py"), recursive=True):
        called |= set(re.findall(r"lib\(\)\.(stx_\w+)", open(path).read()))
    missing = sorted(called - set(_lib._SIGNATURES))
    assert not missing, missing
    lib = _lib.lib()
    with pytest.raises(_lib.HipLibraryError):
        lib.stx_not_a_registered_name


def test_each_module_registers_what_it_calls():
    """Per module, in a fresh interpreter: importing the module alone
    registers every entry point its own source calls."""
    import subprocess
    import sys

    for path in sorted(glob.glob(os.path.join(ROOT, "subsetix_amd", "**", "*.py"), recursive=True)):
        names = sorted(set(re.findall(r"lib\(\)\.(stx_\w+)", open(path).read())))
        if not names:
            continue
        mod = os.path.relpath(path, ROOT)[:-3].replace(os.sep, ".").replace(".__init__", "")
        code = (f"import {mod}\nfrom subsetix_amd import _lib\n"
                f"missing = [n for n in {names!r} if n not in _lib._SIGNATURES]\n"
                "assert not missing, missing\n")
        r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, capture_output=True, text=True)
        assert r.returncode == 0, (mod, r.stderr[-500:])


def test_product_has_no_oracle_dependency():
    for path in glob.glob(os.path.join(ROOT, "subsetix_amd", "**", "*.py"), recursive=True):
        text = open(path).read()
        assert "oracle" not in re.sub(r"#.*", "", text).replace("the oracle", ""), path
