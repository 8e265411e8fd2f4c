"""Seeded synthetic workloads (host NumPy generators).

These build the inputs of the benchmark configurations as plain int32 CSR
arrays on the host; callers move them to the device.  Distributions follow
the reference's generator conventions (subsetix_cupy/benchmarks/utils.py)
restated for the configurations measured here.
"""

from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class HostSet:
    row_offsets: np.ndarray
    begin: np.ndarray
    end: np.ndarray
    rows: np.ndarray

    @property
    def n(self) -> int:
        return int(self.begin.size)


def random_rows_set(rows: int, width: int, *, seed: int, mean_count: float = 32.0, mean_len: float = 40.0,
                    mean_gap: float = 60.0) -> HostSet:
    """Per row: n ~ Poisson(mean_count) intervals; gap and length are
    geometric with the given means (support >= 1, so runs never touch); a
    row stops early at ``width``.  Canonical output (sorted, disjoint)."""
    rng = np.random.default_rng(seed)
    counts = rng.poisson(mean_count, size=rows)
    total = int(counts.sum())
    gaps = rng.geometric(1.0 / mean_gap, size=total).astype(np.int64)
    lens = rng.geometric(1.0 / mean_len, size=total).astype(np.int64)
    off = np.zeros(rows + 1, np.int64)
    off[1:] = np.cumsum(counts)
    # positions: per-row cumulative sums of gap+len
    steps = gaps + lens
    csum = np.cumsum(steps)
    row_start = np.repeat(off[:-1], counts)
    base = np.where(row_start > 0, csum[np.maximum(row_start - 1, 0)], 0)
    base = np.where(row_start > 0, base, 0)
    end = csum - base
    begin = end - lens
    keep = end <= width
    # a row stops at the first interval crossing the width
    row_id = np.repeat(np.arange(rows), counts)
    keep_counts = np.bincount(row_id[keep], minlength=rows)
    new_off = np.zeros(rows + 1, np.int64)
    new_off[1:] = np.cumsum(keep_counts)
    return HostSet(new_off.astype(np.int32), begin[keep].astype(np.int32), end[keep].astype(np.int32),
                   np.arange(rows, dtype=np.int32))


def setalgebra_pair(rows: int = 4096, width: int = 65536, seed_a: int = 1, seed_b: int = 2):
    """Set-algebra microbenchmark inputs: two seeded sets on rows x width,
    Poisson(32) intervals per row, geometric lengths (mean 40) and gaps
    (mean 60)."""
    return random_rows_set(rows, width, seed=seed_a), random_rows_set(rows, width, seed=seed_b)


def stack_rows(sets) -> HostSet:
    """Concatenate independent sets row-wise (rows renumbered densely): the
    batched form of many independent instances of a row-local operation."""
    offs, begs, ends = [], [], []
    base = 0
    for s in sets:
        offs.append(s.row_offsets[:-1].astype(np.int64) + base)
        begs.append(s.begin)
        ends.append(s.end)
        base += s.n
    off = np.concatenate(offs + [np.array([base])]).astype(np.int32)
    nrows = off.size - 1
    return HostSet(off, np.concatenate(begs), np.concatenate(ends), np.arange(nrows, dtype=np.int32))


def stack_rows_gapped(sets, gap: int) -> HostSet:
    """Concatenate independent sets row-wise keeping their row ids, each
    set's rows shifted past the previous set's last row plus ``gap`` empty
    rows: the batched form of a row-window operation (dilation by up to
    ``gap`` / 2 rows keeps the instances apart)."""
    offs, begs, ends, rows = [], [], [], []
    base, shift = 0, 0
    for s in sets:
        offs.append(s.row_offsets[:-1].astype(np.int64) + base)
        begs.append(s.begin)
        ends.append(s.end)
        rows.append(s.rows.astype(np.int64) + shift)
        base += s.n
        shift += (int(s.rows[-1]) + 1 if s.rows.size else 0) + gap
    off = np.concatenate(offs + [np.array([base])]).astype(np.int32)
    return HostSet(off, np.concatenate(begs), np.concatenate(ends), np.concatenate(rows).astype(np.int32))


# ---------------------------------------------------------------------------
# Shapes -> interval sets (host, vectorised)

def _merge_segments(rows: np.ndarray, begin: np.ndarray, end: np.ndarray, nrows: int, width: int) -> HostSet:
    """Union of (row, [b, e)) segments: maximal runs per row, rows 0..nrows-1."""
    keep = end > begin
    rows, begin, end = rows[keep], begin[keep], end[keep]
    order = np.lexsort((begin, rows))
    rows, begin, end = rows[order], begin[order], end[order]
    stride = np.int64(width) + 8
    gb = rows.astype(np.int64) * stride + begin
    ge = rows.astype(np.int64) * stride + end
    run_max = np.maximum.accumulate(ge)
    prev = np.concatenate([[np.iinfo(np.int64).min], run_max[:-1]])
    start = gb > prev
    idx = np.nonzero(start)[0]
    out_b = begin[idx]
    out_r = rows[idx]
    # run end = running max just before the next start
    nxt = np.concatenate([idx[1:], [gb.size]])
    out_e = (run_max[nxt - 1] - out_r.astype(np.int64) * stride).astype(np.int64)
    counts = np.bincount(out_r, minlength=nrows)
    off = np.zeros(nrows + 1, np.int64)
    off[1:] = np.cumsum(counts)
    return HostSet(off.astype(np.int32), out_b.astype(np.int32), out_e.astype(np.int32),
                   np.arange(nrows, dtype=np.int32))


def disc_segments(cx, cy, r, n):
    """Row chords of discs (cell centres inside) on an n x n grid."""
    rows, bs, es = [], [], []
    ys = np.arange(n, dtype=np.float64) + 0.5
    for x0, y0, rad in zip(cx, cy, r):
        lo = max(0, int(np.floor(y0 - rad)))
        hi = min(n, int(np.ceil(y0 + rad)) + 1)
        yy = ys[lo:hi]
        h2 = rad * rad - (yy - y0) ** 2
        ok = h2 >= 0
        half = np.sqrt(np.where(ok, h2, 0.0))
        b = np.ceil(x0 - half - 0.5).astype(np.int64)
        e = np.floor(x0 + half - 0.5).astype(np.int64) + 1
        b = np.clip(b, 0, n)
        e = np.clip(e, 0, n)
        sel = ok & (e > b)
        rows.append(np.arange(lo, hi)[sel])
        bs.append(b[sel])
        es.append(e[sel])
    return np.concatenate(rows), np.concatenate(bs), np.concatenate(es)


def discs_set(n: int, seed: int, scale: float, ndiscs: int = 64, annulus=True) -> HostSet:
    rng = np.random.default_rng(seed)
    cx = rng.uniform(0, n, ndiscs)
    cy = rng.uniform(0, n, ndiscs)
    rad = rng.uniform(50, 800, ndiscs) * scale * (n / 8192.0)
    ax, ay = rng.uniform(0.25 * n, 0.75 * n, 2)
    r_out = rng.uniform(0.08, 0.15) * n * scale
    r_in = 0.6 * r_out
    rows, b, e = disc_segments(cx, cy, rad, n)
    if annulus:
        ro, bo, eo = disc_segments([ax], [ay], [r_out], n)
        ri, bi, ei = disc_segments([ax], [ay], [r_in], n)
        # annulus = outer chord minus inner chord (two pieces per row)
        inner = dict(zip(ri.tolist(), zip(bi.tolist(), ei.tolist())))
        ar, ab, ae = [], [], []
        for y, lo, hi in zip(ro.tolist(), bo.tolist(), eo.tolist()):
            if y in inner:
                il, ih = inner[y]
                ar += [y, y]
                ab += [lo, ih]
                ae += [il, hi]
            else:
                ar.append(y)
                ab.append(lo)
                ae.append(hi)
        rows = np.concatenate([rows, np.array(ar, np.int64)])
        b = np.concatenate([b, np.array(ab, np.int64)])
        e = np.concatenate([e, np.array(ae, np.int64)])
    return _merge_segments(rows.astype(np.int64), b, e, n, n)


def coverage(s: HostSet, n: int) -> float:
    return float((s.end.astype(np.int64) - s.begin).sum()) / float(n) / float(n)


def stencil_discs_config(n: int = 8192, seed: int = 3, target: float = 0.30, tol: float = 0.01):
    """Large stencil configuration: fine n x n domain covered ~30% by 64
    seeded discs (radii U(50, 800) cells at n = 8192) plus one annulus.  A
    radius scale s is found by bisection so the coverage is target +- tol;
    returns (disc set, s, coverage)."""
    lo, hi = 0.2, 1.5
    best = None
    for _ in range(30):
        mid = 0.5 * (lo + hi)
        s = discs_set(n, seed, mid)
        c = coverage(s, n)
        best = (s, mid, c)
        if abs(c - target) <= tol:
            break
        if c > target:
            hi = mid
        else:
            lo = mid
    return best


def gaussian_bump(n: int, amp=1.0, sigma=0.1, center=(0.3, 0.3)) -> np.ndarray:
    """u0 on cell centres (i + 0.5)/n of the unit square, row-major (y, x)."""
    xs = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xs, xs)
    return amp * np.exp(-((X - center[0]) ** 2 + (Y - center[1]) ** 2) / (2 * sigma ** 2))


def batched_instances(count: int = 512, seed: int = 4, nmin: int = 64, nmax: int = 2048):
    """Batched AMR instances: coarse size N = floor(exp(U(ln nmin, ln nmax))),
    1-8 Gaussian bumps (centre U(0.15, 0.85)^2, sigma U(0.02, 0.12), amplitude
    U(0.5, 1.5)), velocity U(-1, 1)^2; one SeedSequence child per instance."""
    children = np.random.SeedSequence(seed).spawn(count)
    out = []
    for ss in children:
        rng = np.random.default_rng(ss)
        n = int(np.floor(np.exp(rng.uniform(np.log(nmin), np.log(nmax)))))
        k = int(rng.integers(1, 9))
        bumps = [(rng.uniform(0.15, 0.85), rng.uniform(0.15, 0.85), rng.uniform(0.02, 0.12), rng.uniform(0.5, 1.5))
                 for _ in range(k)]
        vel = (float(rng.uniform(-1, 1)), float(rng.uniform(-1, 1)))
        out.append({"n": n, "bumps": bumps, "velocity": vel})
    return out


def bumps_field(n: int, bumps, dtype=np.float32) -> np.ndarray:
    xs = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xs, xs)
    u = np.zeros((n, n), np.float64)
    for cx, cy, sg, amp in bumps:
        u += amp * np.exp(-((X - cx) ** 2 + (Y - cy) ** 2) / (2 * sg * sg))
    return u.astype(dtype)
