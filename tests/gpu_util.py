"""Helpers for the GPU parity tests (product <-> oracle conversions)."""

from __future__ import annotations

import numpy as np

from oracle import oracle as ox


def to_device(s: ox.OSet):
    import subsetix_amd as sx

    return sx.build_interval_set(s.off, s.beg, s.end, rows=s.rows)


def to_host(s) -> ox.OSet:
    return ox.OSet(s.rows.cpu().numpy().astype(np.int32), s.row_offsets.cpu().numpy().astype(np.int32),
                   s.begin.cpu().numpy().astype(np.int32), s.end.cpu().numpy().astype(np.int32))


def hostset_to_oset(h) -> ox.OSet:
    return ox.OSet(h.rows, h.row_offsets, h.begin, h.end)
