"""Schedule autotuner for the kernels that have more than one schedule.

Every candidate of a family computes bit-identical results (each schedule is
parity-tested on its own), so the choice changes time only.  When tuning is
enabled, the first call in a shape bucket runs every candidate on the
caller's own data, times it with HIP events on the current stream, and keeps
the fastest; later calls in the bucket (tuning enabled or not, graph capture
included) take the choice from the table.  A bucket is the family plus
half-octaves of the sizes that decide the schedule (cells or intervals, and
their mean per row), so one tuning serves a regrid cadence or a stream of
similar queries.

Families and candidates:
  setop      flat tiles 1024 x 8, register-output form (-7, two operands
             read as stored), the same with count-balanced tiles from a
             tile table (-8: inputs of uneven density), flat tiles
             1024 x 8 (-6) and 512 x 8 (-5),
             tile kernel 512 x 8 single pass (-1) or two passes (-2),
             128 x 16 (-3), 64 x 32 (-4) (threads x positions per thread),
             one lane group of 4 / 16 / 64 lanes per row (4 / 16 / 64);
             operands with row maps have the lane-group candidates only
  upwind     segment tiles ("seg") and lane groups of 4 / 16 / 64 per row
  transfer   segment tables ("seg", ratio 2) and lane groups per target row
  grad_tag   lane groups of 4 / 16 / 64 per row

A key missing from the table while tuning is off (or inside a graph
capture) takes the choice of the nearest tuned bucket of the same family and
discrete parameters (within two octaves in total), so a geometry that drifts
between regrids keeps the schedule tuned on its first geometry.

The table is per process; ``save`` / ``load`` keep it in a JSON file keyed
by the device name, so a serving process can start from earlier tunings.
"""

from __future__ import annotations

import contextlib
import json
import math
import statistics
from dataclasses import dataclass, field
from typing import Callable, Dict, Hashable, Iterator, Optional, Tuple

import torch

Candidates = Dict[object, Callable[[], None]]


def bucket(x: float) -> float:
    """Half-octave bucket of a positive size (0 for sizes below 1)."""
    return 0.0 if x < 1 else round(2.0 * math.log2(x)) / 2.0


@dataclass
class Entry:
    choice: object
    times_us: Dict[str, float] = field(default_factory=dict)


class Autotuner:
    """Table of (family, bucket...) -> fastest candidate."""

    def __init__(self, reps: int = 3, warmup: int = 1, timer: Optional[Callable[..., float]] = None):
        self.reps = int(reps)
        self.warmup = int(warmup)
        self.enabled = False
        self.table: Dict[Tuple, Entry] = {}
        self._timer = timer or _event_timer
        self._warm = set()  # (family, candidate) pairs launched once already

    def lookup(self, key: Tuple) -> Optional[object]:
        e = self.table.get(key)
        return None if e is None else e.choice

    def nearest(self, key: Tuple, max_dist: float = 2.0) -> Optional[object]:
        """Choice of the closest tuned key that agrees with ``key`` on every
        non-float component (distance: summed octave differences)."""
        best, dist = None, max_dist
        for k, e in self.table.items():
            if len(k) != len(key):
                continue
            d = 0.0
            for x, y in zip(k, key):
                if isinstance(x, float) and isinstance(y, float):
                    d += abs(x - y)
                elif x != y:
                    d = math.inf
                    break
            if d <= dist:
                best, dist = e.choice, d
        return best

    def tune(self, key: Tuple, factory: Callable[[], Candidates]) -> object:
        """Time every candidate and record the fastest (ties keep the first
        in the factory's order).  A candidate's first launch in the process
        is preceded by warm-up runs; one clearly slower than the best so far
        stops after its first timed run."""
        cands = factory()
        if not cands:
            raise ValueError(f"no candidates for {key}")
        times = {}
        best_t = math.inf
        for name, fn in cands.items():
            tag = (key[0], name)
            if tag not in self._warm:
                for _ in range(self.warmup):
                    fn()
                self._warm.add(tag)
            times[name] = self._timer(fn, self.reps, 2.0 * best_t)
            best_t = min(best_t, times[name])
        best = min(cands, key=lambda n: times[n])
        self.table[key] = Entry(best, {str(k): v for k, v in times.items()})
        return best

    def pick(self, key: Tuple, factory: Callable[[], Candidates]) -> Optional[object]:
        """The table's choice for ``key``; else, when enabled and the current
        stream is not capturing a graph, tune now; else the nearest tuned
        bucket's choice, or None (the caller's fixed rule applies)."""
        hit = self.lookup(key)
        if hit is not None:
            return hit
        if not self.enabled or _capturing():
            return self.nearest(key)
        return self.tune(key, factory)

    def summary(self) -> Dict[str, dict]:
        return {"/".join(str(k) for k in key): {"choice": str(e.choice), "us": e.times_us}
                for key, e in sorted(self.table.items(), key=lambda kv: str(kv[0]))}

    def save(self, path: str, device=None) -> None:
        name = _device_name(device)
        data = {}
        try:
            with open(path) as f:
                data = json.load(f)
        except (OSError, ValueError):
            data = {}
        data[name] = [{"key": list(k), "choice": e.choice, "us": e.times_us} for k, e in self.table.items()]
        with open(path, "w") as f:
            json.dump(data, f, indent=1)

    def load(self, path: str, device=None) -> int:
        """Merge the entries saved for this device; returns how many."""
        with open(path) as f:
            data = json.load(f)
        n = 0
        for item in data.get(_device_name(device), []):
            self.table[tuple(item["key"])] = Entry(item["choice"], dict(item.get("us", {})))
            n += 1
        return n


def _capturing() -> bool:
    return torch.cuda.is_available() and torch.cuda.is_current_stream_capturing()


def _device_name(device=None) -> str:
    if not torch.cuda.is_available():
        return "cpu"
    return torch.cuda.get_device_name(device)


def _event_timer(fn: Callable[[], None], reps: int, bound: float = math.inf) -> float:
    """Median microseconds per call over ``reps`` timed calls; a first call
    slower than ``bound`` is returned alone."""
    out = []
    for _ in range(max(1, reps)):
        start, stop = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
        fn()
        stop.record()
        stop.synchronize()
        out.append(start.elapsed_time(stop) * 1e3)
        if out[0] > bound:
            break
    return float(statistics.median(out))


_DEFAULT = Autotuner()


def default() -> Autotuner:
    return _DEFAULT


def enable(on: bool = True) -> None:
    _DEFAULT.enabled = bool(on)


def enabled() -> bool:
    return _DEFAULT.enabled


@contextlib.contextmanager
def tuning(on: bool = True) -> Iterator[Autotuner]:
    """Enable (or disable) tuning inside the block."""
    prev = _DEFAULT.enabled
    _DEFAULT.enabled = bool(on)
    try:
        yield _DEFAULT
    finally:
        _DEFAULT.enabled = prev


def pick(key: Tuple[Hashable, ...], factory: Callable[[], Candidates]) -> Optional[object]:
    return _DEFAULT.pick(key, factory)


def reset() -> None:
    _DEFAULT.table.clear()


GROUPS = (4, 16, 64)
