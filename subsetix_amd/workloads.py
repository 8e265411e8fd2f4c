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
