"""ctypes binding of libsubsetix_hip.so and the per-stream device context.

The library is the product path: every interval / field operation in this
package runs one of its gfx950 kernels.  There is no CPU fallback: importing
an operation without the built library, or calling one without a HIP device,
raises.
"""

from __future__ import annotations

import ctypes
import os
import threading
from typing import Dict, Tuple

import torch

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsubsetix_hip.so")

_P = ctypes.c_void_p
_I32 = ctypes.c_int32
_I64 = ctypes.c_int64
_F32 = ctypes.c_float
_F64 = ctypes.c_double

# name -> (restype, argtypes)
_SIGNATURES: Dict[str, Tuple[object, list]] = {
    "stx_abi_version": (_I32, []),
    "stx_last_error": (ctypes.c_char_p, []),
    "stx_set_spin_limit": (None, [_I64]),
    "stx_scan_i32": (_I32, [_P, _P, _I64, _P, _P, _P, _P, _I64, _I32, _P]),
    "stx_scan_scratch_bytes": (_I64, [_I64]),
    "stx_setop_scratch_bytes": (_I64, [_I64, _I32]),
    "stx_setop_scratch_bytes_n": (_I64, [_I64, _I64, _I32]),
    "stx_setop": (_I32, [_I32, _P, _P, _P, _P, _I32, _I32, _I64, _P, _P, _P, _P, _I32, _I32, _I64, _P, _I64, _I32,
                         _P, _P, _P, _I64, _P, _P, _P, _I64, _I32, _P]),
    "stx_rows_union": (_I32, [_P, _I64, _I32, _P, _I64, _I32, _P, _P, _P, _P, _P, _P, _P, _I64, _I32, _P]),
    "stx_row_instance": (_I32, [_P, _I64, _P, _I32, _P, _P]),
    "stx_instance_mass": (_I32, [_I32, _P, _P, _I32, _P, _P, _P]),
}

_lib = None
_handle = None
_lib_lock = threading.Lock()


class HipLibraryError(RuntimeError):
    pass


class _Library:
    """The loaded library, exposing only entry points with a registered
    signature: ctypes would pass the 64-bit device pointers of an
    unprototyped call as 32-bit ints."""

    def __init__(self, handle: ctypes.CDLL):
        self._handle = handle

    def __getattr__(self, name: str):
        if name not in _SIGNATURES:
            raise HipLibraryError(f"{name} has no registered signature (subsetix_amd._lib.register)")
        return getattr(self._handle, name)


def lib() -> _Library:
    """Load libsubsetix_hip.so (built by ``__graft_entry__.build()`` or
    ``make -C subsetix_amd/csrc``).  Raises if it is missing."""
    global _lib, _handle
    if _lib is not None:
        return _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise HipLibraryError(
                    f"{LIB_PATH} is not built; run `make -C subsetix_amd/csrc` (hipcc, gfx950)"
                )
            handle = ctypes.CDLL(LIB_PATH)
            for name, (res, args) in _SIGNATURES.items():
                fn = getattr(handle, name)
                fn.restype = res
                fn.argtypes = args
            _handle = handle
            _lib = _Library(handle)
    return _lib


def register(name: str, restype, argtypes) -> None:
    """Declare the signature of an exported function (used by sibling
    modules for their own entry points)."""
    _SIGNATURES[name] = (restype, argtypes)
    if _handle is not None:
        fn = getattr(_handle, name)
        fn.restype = restype
        fn.argtypes = argtypes


def exported_symbols():
    return list(_SIGNATURES)


class StxError(RuntimeError):
    pass


def check(rc: int, what: str = "") -> None:
    if rc != 0:
        msg = lib().stx_last_error()
        msg = msg.decode() if msg else ""
        raise StxError(f"{what or 'subsetix_hip'} failed (status {rc}): {msg}")


def ptr(t) -> int:
    """Device pointer of a tensor (None -> NULL)."""
    if t is None:
        return None
    return t.data_ptr()


# ---------------------------------------------------------------------------
# Per (device, stream) context: status word, look-back scratch, result slots.

STATUS_BITS = {1: "output capacity exceeded", 2: "look-back wait bound exceeded", 4: "alignment violated",
               8: "row ids out of range", 16: "cells outside the domain of the segment table"}


class Context:
    def __init__(self, device: torch.device, stream: torch.cuda.Stream):
        self.device = device
        self.stream_obj = stream
        # [0] status bits, [1..] result slots
        self.words = torch.zeros(64, dtype=torch.int32, device=device)
        self._scratch = torch.zeros(1 << 16, dtype=torch.uint8, device=device)
        self._epoch = 0

    @property
    def stream(self) -> int:
        return self.stream_obj.cuda_stream

    @property
    def status_ptr(self) -> int:
        return self.words.data_ptr()

    def slot(self, k: int) -> int:
        return self.words.data_ptr() + 4 * (1 + k)

    def scratch(self, nbytes: int, count: int = 1) -> Tuple[int, int, int]:
        """Look-back scratch of at least ``nbytes`` and ``count`` fresh
        consecutive launch epochs, the first returned (state words of earlier
        launches then read as unpublished)."""
        if nbytes > self._scratch.numel():
            self._scratch = torch.zeros(max(nbytes, 2 * self._scratch.numel()), dtype=torch.uint8, device=self.device)
        if self._epoch + count >= (1 << 30) - 1:
            self._scratch.zero_()
            self._epoch = 0
        first = self._epoch + 1
        self._epoch += count
        return self._scratch.data_ptr(), self._scratch.numel(), first

    def read_slots(self, n: int):
        """One D2H copy: [status, slot0, slot1, ...]; raises on a device flag."""
        vals = self.words[: n + 1].tolist()
        if vals[0]:
            bits = vals[0]
            self.words[0].zero_()
            msgs = [m for b, m in STATUS_BITS.items() if bits & b]
            raise StxError("device reported: " + ", ".join(msgs))
        return vals[1:]

    def check_status(self) -> None:
        self.read_slots(0)


_contexts: Dict[Tuple[int, int], Context] = {}


def context(device=None) -> Context:
    if not torch.cuda.is_available():
        raise HipLibraryError("subsetix_amd needs a HIP device (torch.cuda.is_available() is False)")
    lib()
    dev = torch.device("cuda", torch.cuda.current_device()) if device is None else torch.device(device)
    stream = torch.cuda.current_stream(dev)
    key = (dev.index if dev.index is not None else 0, stream.cuda_stream)
    ctx = _contexts.get(key)
    if ctx is None:
        ctx = Context(dev, stream)
        _contexts[key] = ctx
    return ctx


def default_device() -> torch.device:
    return torch.device("cuda", torch.cuda.current_device())
