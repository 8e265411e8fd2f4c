"""Micro-benchmark suite of the interval-set path (drop-in for
``subsetix_cupy.benchmarks``: the case names, geometries and ns-per-input-
interval report of subsetix_cupy/benchmarks/__init__.py, cases.py:549-846
and __main__.py:71-86), timed on the HIP device.

``run_case`` times each call with a pair of HIP events on the current stream
(GPU time, including any host wait inside the call) and the host clock (CPU
time), after ``warmup`` untimed calls.  ``python -m subsetix_amd.benchmarks``
lists or runs the cases.
"""

from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Any, Callable, Dict, List, Optional

import numpy as np


@dataclass
class BenchmarkTarget:
    func: Callable[[], Any]
    repeat: int = 50
    warmup: int = 5
    metadata: Dict[str, Any] | None = None


@dataclass
class BenchmarkCase:
    name: str
    description: str
    setup: Callable[[Any], BenchmarkTarget]


@dataclass
class BenchmarkResult:
    gpu_times: np.ndarray = field(default_factory=lambda: np.zeros(0))  # seconds per call
    cpu_times: np.ndarray = field(default_factory=lambda: np.zeros(0))


def run_case(case: BenchmarkCase, *, repeat: Optional[int] = None, warmup: Optional[int] = None):
    """(result, metadata, repeat, warmup) of one case."""
    import torch

    target = case.setup(torch)
    reps = target.repeat if repeat is None else int(repeat)
    warm = target.warmup if warmup is None else int(warmup)
    if reps <= 0:
        raise ValueError("repeat has to be at least 1")
    if warm < 0:
        raise ValueError("warmup cannot be negative")
    for _ in range(warm):
        target.func()
    torch.cuda.synchronize()
    starts = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    stops = [torch.cuda.Event(enable_timing=True) for _ in range(reps)]
    cpu = np.zeros(reps)
    for k in range(reps):
        t0 = time.perf_counter()
        starts[k].record()
        target.func()
        stops[k].record()
        cpu[k] = time.perf_counter() - t0
    torch.cuda.synchronize()
    gpu = np.array([a.elapsed_time(b) * 1e-3 for a, b in zip(starts, stops)])
    return BenchmarkResult(gpu, cpu), target.metadata or {}, reps, warm


def load_cases() -> List[BenchmarkCase]:
    from .cases import CASES

    return list(CASES)


__all__ = ["BenchmarkCase", "BenchmarkTarget", "BenchmarkResult", "load_cases", "run_case"]
