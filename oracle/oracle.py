"""CPU oracle for the subsetix interval-set hot path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module, and only as the checker.  The product package
``subsetix_amd`` never imports it.

The primitives live in ``oracle.c`` (built into ``liboracle.so`` by the
Makefile next to this file); this module wraps them with ctypes and restates
the reference's Python-level compositions (erode, ghost zones, grading,
field synchronisation, geometry) on top of them, citing the reference file and
line each one follows (paths relative to the sbstndb/subsetix repository).

Sets are plain ``OSet(rows, off, beg, end)`` tuples of int32 NumPy arrays.
"""

from __future__ import annotations

import ctypes
import os
import subprocess
from typing import NamedTuple, Optional

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liboracle.so")


class OSet(NamedTuple):
    rows: np.ndarray
    off: np.ndarray
    beg: np.ndarray
    end: np.ndarray

    @property
    def n(self) -> int:
        return int(self.beg.size)

    @property
    def nrows(self) -> int:
        return int(self.rows.size)


class _CSet(ctypes.Structure):
    _fields_ = [
        ("nrows", ctypes.c_int64),
        ("n", ctypes.c_int64),
        ("rows", ctypes.POINTER(ctypes.c_int32)),
        ("off", ctypes.POINTER(ctypes.c_int32)),
        ("beg", ctypes.POINTER(ctypes.c_int32)),
        ("end", ctypes.POINTER(ctypes.c_int32)),
    ]


_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.ox_mass_f32.restype = ctypes.c_double
        _lib.ox_mass_f64.restype = ctypes.c_double
    return _lib


def _i32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a, dtype=np.int32))


def make(row_offsets, begin, end, rows=None) -> OSet:
    off = _i32(row_offsets)
    if rows is None:
        rows = np.arange(off.size - 1, dtype=np.int32)
    return OSet(_i32(rows), off, _i32(begin), _i32(end))


def _ptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_int32))


def _to_c(s: OSet) -> _CSet:
    c = _CSet()
    c.nrows = s.rows.size
    c.n = s.beg.size
    c.rows = _ptr(s.rows)
    c.off = _ptr(s.off)
    c.beg = _ptr(s.beg)
    c.end = _ptr(s.end)
    c._keep = s  # keep arrays alive
    return c


def _from_c(c: _CSet) -> OSet:
    nr, n = int(c.nrows), int(c.n)
    rows = np.ctypeslib.as_array(c.rows, shape=(max(nr, 1),))[:nr].copy()
    off = np.ctypeslib.as_array(c.off, shape=(nr + 1,)).copy()
    beg = np.ctypeslib.as_array(c.beg, shape=(max(n, 1),))[:n].copy()
    end = np.ctypeslib.as_array(c.end, shape=(max(n, 1),))[:n].copy()
    lib().ox_free(ctypes.byref(c))
    return OSet(rows, off, beg, end)


class OracleError(ValueError):
    pass


def _check(rc: int) -> None:
    if rc != 0:
        raise OracleError(f"oracle error {rc}")


OPS = {"union": 0, "intersection": 1, "difference": 2}


def setop(op: str, a: OSet, b: OSet) -> OSet:
    """expressions.py:446-528 + kernels.py:11-236."""
    out = _CSet()
    _check(lib().ox_setop(OPS[op], ctypes.byref(_to_c(a)), ctypes.byref(_to_c(b)), ctypes.byref(out)))
    return _from_c(out)


def union(a, b):
    return setop("union", a, b)


def intersection(a, b):
    return setop("intersection", a, b)


def difference(a, b):
    return setop("difference", a, b)


def symmetric_difference(a, b):
    """expressions.py:380-388: (A-B) U (B-A)."""
    return union(difference(a, b), difference(b, a))


def interval_rows(s: OSet) -> np.ndarray:
    """expressions.py:236-249."""
    return np.repeat(s.rows, np.diff(s.off)).astype(np.int32)


def rebuild(rows, beg, end, rows_hint=None) -> OSet:
    """morphology.py:21-107 (_build_interval_set_from_rows)."""
    rows, beg, end = _i32(rows), _i32(beg), _i32(end)
    hint = _i32(rows_hint if rows_hint is not None else [])
    out = _CSet()
    _check(lib().ox_rebuild(_ptr(rows), _ptr(beg), _ptr(end), ctypes.c_int64(beg.size), _ptr(hint),
                            ctypes.c_int64(hint.size), ctypes.c_int(rows_hint is not None), ctypes.byref(out)))
    return _from_c(out)


def translate(s: OSet, dx: int = 0, dy: int = 0) -> OSet:
    """morphology.py:115-140."""
    out = _CSet()
    _check(lib().ox_translate(ctypes.byref(_to_c(s)), ctypes.c_int32(dx), ctypes.c_int32(dy), ctypes.byref(out)))
    return _from_c(out)


def dilate(s: OSet, hx: int = 1, hy: int = 1) -> OSet:
    """morphology.py:211-265."""
    out = _CSet()
    _check(lib().ox_dilate(ctypes.byref(_to_c(s)), ctypes.c_int32(hx), ctypes.c_int32(hy), ctypes.byref(out)))
    return _from_c(out)


def full(width: int, height: int) -> OSet:
    """morphology.py:344-358."""
    out = _CSet()
    _check(lib().ox_full(ctypes.c_int32(width), ctypes.c_int32(height), ctypes.byref(out)))
    return _from_c(out)


def clip(s: OSet, width: int, height: int) -> OSet:
    """morphology.py:268-319."""
    out = _CSet()
    _check(lib().ox_clip(ctypes.byref(_to_c(s)), ctypes.c_int32(width), ctypes.c_int32(height), ctypes.byref(out)))
    return _from_c(out)


def ghost_zones(s: OSet, hx: int = 1, hy: int = 1, *, width: int, height: Optional[int] = None) -> OSet:
    """morphology.py:322-341."""
    if height is None:
        height = s.off.size - 1
    dil = clip(dilate(s, hx, hy), width, height)
    base = clip(s, width, height)
    return difference(dil, base)


def erode(s: OSet, hx: int = 1, hy: int = 1, *, width: int, height: Optional[int] = None) -> OSet:
    """morphology.py:361-386: U - clip(dilate(U - clip(S)))."""
    if height is None:
        height = s.off.size - 1
    if width <= 0 or height < 0:
        raise OracleError("bad domain")
    base = clip(s, width, height)
    uni = full(width, height)
    comp = difference(uni, base)
    dil = clip(dilate(comp, hx, hy), width, height)
    return difference(uni, dil)


def prolong_set(s: OSet, ratio: int, with_base: bool = False):
    """multilevel.py:213-269."""
    out = _CSet()
    base = ctypes.POINTER(ctypes.c_int32)()
    _check(lib().ox_prolong_set(ctypes.byref(_to_c(s)), ctypes.c_int32(ratio), ctypes.byref(out), ctypes.byref(base)))
    res = _from_c(out)
    nb = res.n if ratio > 1 or True else 0
    b = np.ctypeslib.as_array(base, shape=(max(nb, 1),))[:nb].copy()
    lib().ox_free_buf(base)
    return (res, b) if with_base else res


def restrict_set(s: OSet, ratio: int) -> OSet:
    """multilevel.py:272-362."""
    out = _CSet()
    _check(lib().ox_restrict_set(ctypes.byref(_to_c(s)), ctypes.c_int32(ratio), ctypes.byref(out)))
    return _from_c(out)


def cell_offsets(s: OSet) -> np.ndarray:
    """interval_field.py:64-75."""
    o = np.zeros(s.n + 1, dtype=np.int64)
    np.cumsum((s.end - s.beg).astype(np.int64), out=o[1:])
    return o.astype(np.int32)


def prolong_field(c: OSet, cvals: np.ndarray, ratio: int, dtype=np.float32):
    """multilevel.py:365-407 + kernels.py:384-455 (prolong_set, then copy)."""
    f = prolong_set(c, ratio)
    cv = np.ascontiguousarray(np.asarray(cvals, dtype=dtype))
    fv = np.zeros(int(cell_offsets(f)[-1]), dtype=dtype)
    if fv.size == 0 or cv.size == 0:
        return f, fv
    if ratio == 1:
        return f, cv.copy()
    _check(lib().ox_prolong_field(ctypes.byref(_to_c(c)), cv.ctypes.data_as(ctypes.c_void_p), ctypes.byref(_to_c(f)),
                                  fv.ctypes.data_as(ctypes.c_void_p), ctypes.c_int32(ratio), ctypes.c_int(cv.itemsize)))
    return f, fv


REDUCERS = {"mean": 0, "sum": 1, "min": 2, "max": 3}


def restrict_field(f: OSet, fvals: np.ndarray, ratio: int, reducer: str = "mean", accumulate: str = "f32"):
    """multilevel.py:410-479 + kernels.py:458-598.

    accumulate="f32": the reference FP32 kernel (float in, float accumulate).
    accumulate="f64": float in, double accumulate, float out.
    fvals of dtype float64: double throughout.
    """
    c = restrict_set(f, ratio)
    fv = np.ascontiguousarray(fvals)
    if fv.dtype == np.float64:
        kind, odt = 2, np.float64
    else:
        fv = np.ascontiguousarray(fv, dtype=np.float32)
        kind, odt = (1 if accumulate == "f64" else 0), np.float32
    cv = np.zeros(int(cell_offsets(c)[-1]), dtype=odt)
    if fv.size == 0 or cv.size == 0:
        return c, cv
    if ratio == 1:
        return c, fv.astype(odt)
    _check(lib().ox_restrict_field(ctypes.byref(_to_c(c)), cv.ctypes.data_as(ctypes.c_void_p), ctypes.byref(_to_c(f)),
                                   fv.ctypes.data_as(ctypes.c_void_p), ctypes.c_int32(ratio),
                                   ctypes.c_int(REDUCERS[reducer]), ctypes.c_int(kind)))
    return c, cv


def upwind(s: OSet, u: np.ndarray, *, width, height, a, b, dt, dx, dy) -> np.ndarray:
    """interval_stencil.py:16-219.  FP32 when u is float32, else FP64."""
    u = np.ascontiguousarray(u)
    out = np.zeros_like(u)
    if s.n == 0:
        return out
    if u.dtype == np.float32:
        f, ct = lib().ox_upwind_f32, ctypes.c_float
    else:
        u = np.ascontiguousarray(u, dtype=np.float64)
        out = np.zeros_like(u)
        f, ct = lib().ox_upwind_f64, ctypes.c_double
    _check(f(ctypes.byref(_to_c(s)), u.ctypes.data_as(ctypes.c_void_p), out.ctypes.data_as(ctypes.c_void_p),
             ctypes.c_int32(width), ctypes.c_int32(height), ct(a), ct(b), ct(dt), ct(dx), ct(dy)))
    return out


def mass(u: np.ndarray, cell_area: float) -> float:
    u = np.ascontiguousarray(u)
    if u.dtype == np.float32:
        return float(lib().ox_mass_f32(u.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(u.size), ctypes.c_double(cell_area)))
    u = np.ascontiguousarray(u, dtype=np.float64)
    return float(lib().ox_mass_f64(u.ctypes.data_as(ctypes.c_void_p), ctypes.c_int64(u.size), ctypes.c_double(cell_area)))


def gradient(data: np.ndarray, width: int, height: int) -> np.ndarray:
    """regrid.py:175-274 (dense W x H FP32)."""
    d = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
    out = np.empty_like(d)
    lib().ox_gradient_f32(d.ctypes.data_as(ctypes.c_void_p), ctypes.c_int32(width), ctypes.c_int32(height),
                          out.ctypes.data_as(ctypes.c_void_p))
    return out


def threshold(data: np.ndarray, width: int, height: int, t: float) -> OSet:
    """regrid.py:277-333: maximal runs of v >= t, rows arange(H)."""
    d = np.ascontiguousarray(data, dtype=np.float32).reshape(-1)
    out = _CSet()
    _check(lib().ox_threshold(d.ctypes.data_as(ctypes.c_void_p), ctypes.c_int32(width), ctypes.c_int32(height),
                              ctypes.c_float(t), ctypes.byref(out)))
    return _from_c(out)


def gradient_tag(data, width, height, threshold_value, epsilon=1e-8) -> OSet:
    """regrid.py:336-350."""
    g = gradient(data, width, height)
    return threshold(g, width, height, max(float(threshold_value), float(epsilon)))


def grading(s: OSet, *, padding=1, mode="von_neumann", width, height) -> OSet:
    """regrid.py:404-442."""
    if padding <= 0:
        return s
    if mode not in ("von_neumann", "moore"):
        raise OracleError("mode")
    if mode == "moore":
        exp = dilate(s, padding, padding)
    else:
        exp = union(dilate(s, padding, 0), dilate(s, 0, padding))
    exp = clip(exp, width, height)
    base = clip(s, width, height)
    return clip(union(base, exp), width, height)


def subset_copy(mode: str, sup: OSet, sup_vals: np.ndarray, sub: OSet, sub_vals: np.ndarray) -> None:
    """fields.py:127-294; mode 'gather' (super -> subset) or 'scatter'."""
    assert sup_vals.flags.c_contiguous and sub_vals.flags.c_contiguous
    _check(lib().ox_subset_copy(ctypes.c_int(0 if mode == "gather" else 1), ctypes.byref(_to_c(sup)),
                                sup_vals.ctypes.data_as(ctypes.c_void_p), ctypes.byref(_to_c(sub)),
                                sub_vals.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(sup_vals.itemsize)))


def gather(sup: OSet, sup_vals, sub: OSet):
    vals = np.zeros(int(cell_offsets(sub)[-1]), dtype=sup_vals.dtype)
    subset_copy("gather", sup, np.ascontiguousarray(sup_vals), sub, vals)
    return vals


def compact_rows(s: OSet) -> OSet:
    """Drop the rows that hold no interval (cells are unchanged)."""
    keep = np.diff(s.off) > 0
    off = np.concatenate([[0], np.cumsum(np.diff(s.off)[keep])]).astype(np.int32)
    return OSet(s.rows[keep].copy(), off, s.beg.copy(), s.end.copy())


def synchronize(coarse: OSet, cvals, fine: OSet, fvals, refine: OSet, *, ratio, accumulate="f32",
                fill_fine_outside=True):
    """fields.py:570-612 (copy=True form): restrict fine onto refine, then
    prolong coarse_only onto the fine level.  Returns new (cvals, fvals).

    coarse_only is row-compacted before the prolongation: the reference raises
    there when a coarse row is fully refined (prolong_set drops the empty row,
    multilevel.py:255, and the row check at multilevel.py:184-185 fails); the
    compacted form gives the cell-wise result the reference computes whenever
    it does not raise."""
    cv = np.array(cvals, copy=True)
    fv = np.array(fvals, copy=True)
    rset, rvals = restrict_field(fine, fvals, ratio, "mean", accumulate)
    if refine.n:
        sub = gather(rset, rvals.astype(cv.dtype), refine)
        subset_copy("scatter", coarse, cv, refine, sub)
    if fill_fine_outside:
        co = compact_rows(difference(coarse, refine))
        if co.n:
            csub = gather(coarse, cv, co)
            fset, fsub = prolong_field(co, csub, ratio, dtype=cv.dtype)
            if fsub.size:
                subset_copy("scatter", fine, fv, fset, fsub)
    return cv, fv


_KIND = {(np.dtype(np.float32), "f32"): 0, (np.dtype(np.float32), "f64"): 1}


def _transfer(mode: int, target: OSet, dst: OSet, dvals: np.ndarray, src: OSet, svals: np.ndarray, ratio: int,
              accumulate: str) -> None:
    assert dvals.flags.c_contiguous and svals.flags.c_contiguous and dvals.dtype == svals.dtype
    kind = 2 if dvals.dtype == np.float64 else _KIND[(dvals.dtype, accumulate)]
    _check(lib().ox_transfer(ctypes.c_int(mode), ctypes.c_int(kind), ctypes.c_int32(ratio),
                             ctypes.byref(_to_c(target)), ctypes.byref(_to_c(dst)),
                             dvals.ctypes.data_as(ctypes.c_void_p), ctypes.byref(_to_c(src)),
                             svals.ctypes.data_as(ctypes.c_void_p)))


def prolong_onto(target: OSet, dst: OSet, dvals: np.ndarray, src: OSet, svals: np.ndarray, ratio: int) -> None:
    """dst(y, x) = src(floor(y/R), floor(x/R)) on the cells of ``target``, in
    place: the prolong + scatter half of fields.py:570-612 (multilevel.py:
    365-407) cell by cell."""
    _transfer(0, target, dst, dvals, src, svals, ratio, "f32")


def restrict_onto(target: OSet, dst: OSet, dvals: np.ndarray, src: OSet, svals: np.ndarray, ratio: int,
                  accumulate: str = "f64") -> None:
    """dst(y, x) = mean of the children src(yR+dy, xR+dx) (dy-major, missing
    children add nothing) on the cells of ``target``, in place: the restrict +
    gather + scatter half of fields.py:570-612 (multilevel.py:410-479)."""
    _transfer(1, target, dst, dvals, src, svals, ratio, accumulate)


def grad_tag(s: OSet, vals: np.ndarray, t: float) -> OSet:
    """Gradient tag on a sparse level (regrid.py:181-218 differences with
    one-sided forms where a neighbour is missing; regrid.py:20-124 runs)."""
    v = np.ascontiguousarray(vals)
    out = _CSet()
    _check(lib().ox_grad_tag(ctypes.byref(_to_c(s)), v.ctypes.data_as(ctypes.c_void_p), ctypes.c_int(v.itemsize),
                             ctypes.c_double(t), ctypes.byref(out)))
    return _from_c(out)


def zeros_on(s: OSet, dtype) -> np.ndarray:
    return np.zeros(int(cell_offsets(s)[-1]), dtype=dtype)


class Composite:
    """Sparse two-level composite step: the reference's step (simulation.py:
    232-354: synchronise, advance both levels, synchronise) on a full coarse
    level and a fine level that covers prolong(refine) plus a one-cell ghost
    ring (morphology.py:322-341).  One step:
      1. prolong coarse -> fine ghost cells;
      2. upwind on the coarse level and on the extended fine level
         (interval_stencil.py:16-219);
      3. restrict fine -> coarse on the refined cells (sums in ``accumulate``).
    """

    def __init__(self, width: int, height: int, refine: OSet, *, velocity=(1.0, 0.5), cfl=0.5, dtype=np.float32,
                 accumulate="f64", ratio=2):
        self.W, self.H, self.R = int(width), int(height), int(ratio)
        self.coarse = full(self.W, self.H)
        self.refine = clip(refine, self.W, self.H)
        self.fine = prolong_set(self.refine, self.R)
        fw, fh = self.W * self.R, self.H * self.R
        self.ghost = ghost_zones(self.fine, 1, 1, width=fw, height=fh)
        self.ext = union(self.fine, self.ghost)
        self.dtype = np.dtype(dtype)
        self.accumulate = accumulate
        self.vel = (float(velocity[0]), float(velocity[1]))
        self.dx_c, self.dy_c = 1.0 / self.W, 1.0 / self.H
        self.dx_f, self.dy_f = self.dx_c / self.R, self.dy_c / self.R
        denom = abs(self.vel[0]) / self.dx_f + abs(self.vel[1]) / self.dy_f
        self.dt = cfl / denom if denom > 0 else 0.0
        self.cv = zeros_on(self.coarse, self.dtype)
        self.fv = zeros_on(self.ext, self.dtype)

    def set_values(self, cv, fv=None) -> None:
        """As the device driver: coarse values, fine values on the extended
        set (default prolonged), then restrict and refresh the ghosts."""
        self.cv = np.ascontiguousarray(cv, dtype=self.dtype).copy()
        if fv is None:
            self.fv = zeros_on(self.ext, self.dtype)
            prolong_onto(self.ext, self.ext, self.fv, self.coarse, self.cv, self.R)
        else:
            self.fv = np.ascontiguousarray(fv, dtype=self.dtype).copy()
        restrict_onto(self.refine, self.coarse, self.cv, self.ext, self.fv, self.R, self.accumulate)
        prolong_onto(self.ghost, self.ext, self.fv, self.coarse, self.cv, self.R)

    def step(self) -> None:
        a, b = self.vel
        prolong_onto(self.ghost, self.ext, self.fv, self.coarse, self.cv, self.R)
        nc = upwind(self.coarse, self.cv, width=self.W, height=self.H, a=a, b=b, dt=self.dt, dx=self.dx_c,
                    dy=self.dy_c)
        nf = upwind(self.ext, self.fv, width=self.W * self.R, height=self.H * self.R, a=a, b=b, dt=self.dt,
                    dx=self.dx_f, dy=self.dy_f)
        restrict_onto(self.refine, self.coarse, nc, self.ext, nf, self.R, self.accumulate)
        self.cv, self.fv = nc, nf

    def mass(self) -> float:
        return mass(self.cv, self.dx_c * self.dy_c)


class Instance3:
    """One instance of the batched three-level advection engine, restated on
    oracle primitives with the instance's rows starting at 0.

    The step follows the reference's step semantics (simulation.py:232-354)
    extended to three levels: prolong ghost cells level by level (coarse to
    fine), upwind every level (interval_stencil.py:16-219, level L with
    dx = dy = 1/(n 2^L)), restrict finest to coarsest onto the refine sets.
    The regrid is the gradient tag (regrid.py:336-350 on the level's cells),
    the grading dilation (regrid.py:404-442, square element), a one-cell
    level-1 ring forced onto its parents and the level-2 tag gated by the
    eroded level 1: the three-level nesting rules of test_amr_grading.py:
    110-248 (level 2 inside level 1, no 8-neighbour contact between level 2
    and cells outside level 1).  New level fields are prolonged from the
    coarser level, then keep the previous values where the level existed.
    """

    def __init__(self, n: int, u0: np.ndarray, velocity, *, dtype=np.float32, levels=3, ratio=2, cfl=0.5,
                 threshold=(0.02, 0.02), grading=1, regrid_every=8, accumulate="f64"):
        self.n, self.L, self.R = int(n), int(levels), int(ratio)
        self.dtype = np.dtype(dtype)
        self.thr = tuple(threshold)
        self.p = int(grading)
        self.k = int(regrid_every)
        self.accumulate = accumulate
        self.D = [full(self.n * self.R ** lv, self.n * self.R ** lv) for lv in range(self.L)]
        finest = self.n * self.R ** (self.L - 1)
        denom = abs(velocity[0]) * finest + abs(velocity[1]) * finest
        dt = cfl / max(denom, 1e-300) if denom > 0 else 0.0
        # coefficients as the device stores them: rounded to the field type
        self.coef = [tuple(float(v) for v in np.array([velocity[0], velocity[1], dt, 1.0 / (self.n * self.R ** lv),
                                                       1.0 / (self.n * self.R ** lv)]).astype(self.dtype))
                     for lv in range(self.L)]
        self.ext = [self.D[0]]
        self.vals = [np.ascontiguousarray(u0, dtype=self.dtype).reshape(-1).copy()]
        self.sets = None
        self.step_count = 0
        self.regrid()
        self.last_regrid = 0

    def _tag(self, s: OSet, v: np.ndarray, lv: int) -> OSet:
        return grad_tag(s, v, self.thr[min(lv, len(self.thr) - 1)])

    def _level_field(self, lv, target, coarser_set, coarser_vals, old_ext, old_vals, old_sets):
        f = zeros_on(target, self.dtype)
        prolong_onto(target, target, f, coarser_set, coarser_vals, self.R)
        if old_sets is not None and lv < len(old_vals) and lv < len(old_sets):
            keep = intersection(target, old_sets[lv])
            prolong_onto(keep, target, f, old_ext[lv], old_vals[lv], 1)
        return f

    def regrid(self) -> None:
        R, p, L, D = self.R, self.p, self.L, self.D
        old_ext, old_vals, old_sets = list(self.ext), list(self.vals), self.sets
        cur0 = self.vals[0]
        tag0 = self._tag(D[0], cur0, 0)
        refine0 = intersection(dilate(tag0, p, p) if p > 0 else tag0, D[0])
        sets, refines = [D[0]], []
        if L >= 2:
            ring = intersection(dilate(prolong_set(refine0, R), 1, 1), D[1])
            refine0 = union(refine0, restrict_set(ring, R))
            l1 = prolong_set(refine0, R)
            sets.append(l1)
            refines.append(refine0)
        if L >= 3:
            tmp1 = self._level_field(1, l1, D[0], cur0, old_ext, old_vals, old_sets)
            comp = difference(D[1], l1)
            eroded = difference(l1, dilate(comp, 1, 1))
            tag1 = self._tag(l1, tmp1, 1)
            refine1 = intersection(dilate(tag1, p, p) if p > 0 else tag1, eroded)
            sets.append(prolong_set(refine1, R))
            refines.append(refine1)
        ghosts, ext, vals = [None], [D[0]], [cur0]
        for lv in range(1, L):
            g = difference(intersection(dilate(sets[lv], 1, 1), D[lv]), sets[lv])
            e = union(sets[lv], g)
            ghosts.append(g)
            ext.append(e)
            vals.append(self._level_field(lv, e, ext[lv - 1], vals[lv - 1], old_ext, old_vals, old_sets))
        self.sets, self.refines, self.ghosts, self.ext, self.vals = sets, refines, ghosts, ext, vals

    def step(self) -> None:
        R, L = self.R, self.L
        for lv in range(1, L):
            prolong_onto(self.ghosts[lv], self.ext[lv], self.vals[lv], self.ext[lv - 1], self.vals[lv - 1], R)
        nxt = []
        for lv in range(L):
            a, b, dt, dx, dy = self.coef[lv]
            w = self.n * R ** lv
            # the level sets may list empty rows just outside the instance
            # (dilation row convention); they hold no cells
            nxt.append(upwind(compact_rows(self.ext[lv]), self.vals[lv], width=w, height=w, a=a, b=b, dt=dt,
                              dx=dx, dy=dy))
        for lv in range(L - 1, 0, -1):
            restrict_onto(self.refines[lv - 1], self.ext[lv - 1], nxt[lv - 1], self.ext[lv], nxt[lv], R,
                          self.accumulate)
        self.vals = nxt

    def run(self, steps: int) -> None:
        for _ in range(int(steps)):
            if self.step_count % self.k == 0 and self.last_regrid != self.step_count:
                self.regrid()
                self.last_regrid = self.step_count
            self.step()
            self.step_count += 1

    def mass(self) -> float:
        return mass(self.vals[0], 1.0 / (self.n * self.n))


def two_level_loop(u0: np.ndarray, n: int, steps: int, regrid_every: int, *, velocity=(1.0, 0.5), cfl=0.5,
                   threshold=0.1, padding=2, mode="moore", tag="value", accumulate="f64", dtype=np.float64):
    """simulation.py:232-354 on oracle primitives: full-domain n x n coarse
    and 2n x 2n fine levels, ratio 2; each step synchronises, regrids every
    ``regrid_every`` steps (value or gradient tag, grading by ``padding``), advances both
    levels and synchronises.  Returns (coarse values, fine values, refine)."""
    w = h = n
    coarse, fine = full(w, h), full(2 * w, 2 * h)
    dxc, dxf = 1.0 / w, 0.5 / w
    dt = cfl / (abs(velocity[0]) / dxf + abs(velocity[1]) / dxf)

    def refine_of(cv):
        d = cv.reshape(h, w).astype(np.float32)
        tagged = gradient_tag(d, w, h, threshold) if tag == "gradient" else threshold_runs(d, w, h, threshold)
        return intersection(grading(tagged, padding=padding, mode=mode, width=w, height=h), coarse)

    cv = np.asarray(u0, dtype=dtype).reshape(-1).copy()
    _, fv = prolong_field(coarse, cv, 2, dtype=dtype)
    refine = refine_of(cv)
    for step in range(steps):
        cv, fv = synchronize(coarse, cv, fine, fv, refine, ratio=2, accumulate=accumulate)
        if step % regrid_every == 0:
            refine = refine_of(cv)
        cv = upwind(coarse, cv, width=w, height=h, a=velocity[0], b=velocity[1], dt=dt, dx=dxc, dy=dxc)
        fv = upwind(fine, fv, width=2 * w, height=2 * h, a=velocity[0], b=velocity[1], dt=dt, dx=dxf, dy=dxf)
        cv, fv = synchronize(coarse, cv, fine, fv, refine, ratio=2, accumulate=accumulate)
    return cv, fv, refine


def threshold_runs(data, width, height, t):
    """regrid.py:277-333 with the max(t, 1e-8) floor of the value tag
    (regrid.py:386-401)."""
    return threshold(data, width, height, max(float(t), 1e-8))


def set_threads(n: int) -> None:
    lib().ox_set_threads(ctypes.c_int(n))


def num_threads() -> int:
    return int(lib().ox_num_threads())


def cells(s: OSet) -> set:
    """Set-of-cells model (test_expressions.py:27-44)."""
    out = set()
    for r in range(s.rows.size):
        row = int(s.rows[r])
        for k in range(int(s.off[r]), int(s.off[r + 1])):
            out.update((row, x) for x in range(int(s.beg[k]), int(s.end[k])))
    return out
