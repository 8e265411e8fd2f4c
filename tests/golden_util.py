"""Helpers shared by the oracle and GPU parity tests: golden fixture loading,
expression-spec evaluation over any backend, bitmap models."""

from __future__ import annotations

import json
import os

import numpy as np

from oracle import oracle as ox

HERE = os.path.dirname(os.path.abspath(__file__))


def load_golden() -> dict:
    with open(os.path.join(HERE, "golden", "reference_vectors.json")) as fh:
        return json.load(fh)


def oset(d) -> ox.OSet:
    return ox.make(d["off"], d["beg"], d["end"], d.get("rows"))


def spans_set(height, spans, rows=None) -> ox.OSet:
    """Dense-row set from {row: [[b, e], ...]} (test_regrid.py:14-29 helper)."""
    off, beg, end = [0], [], []
    row_ids = list(range(height)) if rows is None else list(rows)
    for r in row_ids:
        for b, e in spans.get(str(r), spans.get(r, [])):
            beg.append(b)
            end.append(e)
        off.append(len(beg))
    return ox.make(off, beg, end, row_ids)


def eval_spec(spec, surfaces, ops):
    """Evaluate a nested ["op", lhs, rhs] spec with backend ``ops`` (a dict of
    union/intersection/difference/symmetric_difference/complement)."""
    if isinstance(spec, str):
        return surfaces[spec]
    op, lhs, rhs = spec
    a = eval_spec(lhs, surfaces, ops)
    b = eval_spec(rhs, surfaces, ops)
    if op == "complement":
        return ops["difference"](a, b)
    return ops[op](a, b)


def cell_model(spec, cells):
    if isinstance(spec, str):
        return set(cells[spec])
    op, lhs, rhs = spec
    a = cell_model(lhs, cells)
    b = cell_model(rhs, cells)
    return {"union": a | b, "intersection": a & b, "difference": a - b, "complement": a - b,
            "symmetric_difference": a ^ b}[op]


def assert_set_equal(actual, expected):
    """Bit-exact comparison of rows / row_offsets / begin / end."""
    exp_rows = expected.rows if expected.rows is not None else np.arange(expected.off.size - 1)
    np.testing.assert_array_equal(np.asarray(actual.off, dtype=np.int32), np.asarray(expected.off, dtype=np.int32))
    np.testing.assert_array_equal(np.asarray(actual.beg, dtype=np.int32), np.asarray(expected.beg, dtype=np.int32))
    np.testing.assert_array_equal(np.asarray(actual.end, dtype=np.int32), np.asarray(expected.end, dtype=np.int32))
    np.testing.assert_array_equal(np.asarray(actual.rows, dtype=np.int32), np.asarray(exp_rows, dtype=np.int32))


# ----------------------------------------------------------------------------
# Bitmap models (closed forms) on a bounded window.

def to_bitmap(s: ox.OSet, x0, x1, y0, y1) -> np.ndarray:
    bm = np.zeros((y1 - y0, x1 - x0), dtype=bool)
    for r in range(s.rows.size):
        y = int(s.rows[r])
        for k in range(int(s.off[r]), int(s.off[r + 1])):
            b, e = int(s.beg[k]), int(s.end[k])
            if y0 <= y < y1:
                bm[y - y0, max(b, x0) - x0:max(min(e, x1), x0) - x0] = True
    return bm


def runs_of(bm_row: np.ndarray, x0=0):
    runs = []
    padded = np.concatenate([[False], bm_row, [False]]).astype(np.int8)
    d = np.diff(padded)
    starts = np.nonzero(d == 1)[0]
    stops = np.nonzero(d == -1)[0]
    for b, e in zip(starts, stops):
        runs.append((int(b) + x0, int(e) + x0))
    return runs


def canonical_rows(s: ox.OSet):
    """{row: [(b, e), ...]} of a set (all rows, empty ones included)."""
    out = {}
    for r in range(s.rows.size):
        out[int(s.rows[r])] = [(int(s.beg[k]), int(s.end[k])) for k in range(int(s.off[r]), int(s.off[r + 1]))]
    return out


def random_canonical(rng, nrows, width, mean_n=4, rows=None, touching=False):
    """Random set of maximal runs (non-touching unless ``touching``)."""
    off, beg, end = [0], [], []
    row_ids = np.arange(nrows, dtype=np.int32) if rows is None else np.asarray(rows, dtype=np.int32)
    for _ in range(row_ids.size):
        n = rng.poisson(mean_n)
        x = int(rng.integers(0, 4))
        for _ in range(n):
            gap = int(rng.geometric(0.3)) - (1 if touching and rng.random() < 0.3 else 0)
            b = x + max(gap, 0 if touching else 1)
            e = b + int(rng.geometric(0.25))
            if e > width:
                break
            beg.append(b)
            end.append(e)
            x = e
        off.append(len(beg))
    return ox.make(off, beg, end, row_ids)
