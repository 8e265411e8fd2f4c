"""Served stream: continuous batching of advection jobs and set-algebra
queries on one GPU (no reference counterpart; the reference runs one
simulation or one expression per call, simulation.py:232-354,
expressions.py:446-528).

Requests arrive on a Poisson clock.  The scheduler admits them in arrival
order while their estimated device bytes fit the HBM budget (a request that
does not fit waits; later requests do not overtake it).  The server loop
keeps every admitted request resident and, per iteration:

  * set-algebra queries admitted since the last iteration are stacked row-wise
    into one CSR pair (row ids shifted per query) and answered by ONE fused
    launch per operation (union, intersection, difference); each query's
    result is its slice of the stacked output;
  * advection jobs (two-level composite instances) admitted since the last
    cohort form a COHORT: their instances are stacked into one
    ``StackedComposite`` (one chain of set operations builds the geometry of
    all of them, one segment table per level, four launches per step for the
    whole cohort).  A cohort starts at once when the GPU has none running,
    else when the batching window has passed or it is full; the admitted jobs
    are taken shortest first (by step count), so a cohort's jobs end close
    together.  Cohorts run on
    separate streams and advance in chunks of captured-graph replays; a chunk
    stops at the next step count a job asked for, where that job's fields and
    FP64 mass are copied out.

A job's result does not depend on what else is resident: queries are
row-independent, each cohort instance is bit-identical to the instance run
alone, and every kernel is deterministic.  The plain loop (``plain_loop``)
runs the same requests one at a time with eager launches and a host
synchronisation per step, the reference-shaped form.
"""

from __future__ import annotations

import functools
import math
import time
from collections import deque
from dataclasses import dataclass, field
from typing import Dict, List, Optional

import numpy as np
import torch

from . import _core, workloads
from .expressions import IntervalSet

HBM_BUDGET_BYTES = 200 * (1 << 30)


# ---------------------------------------------------------------- requests
@dataclass
class Request:
    rid: int
    kind: str                 # "advect" or "query"
    arrival: float            # seconds from the start of the run
    payload: dict
    est_bytes: int


@dataclass
class Result:
    rid: int
    kind: str
    arrival: float
    admitted: float
    finished: float
    output: dict = field(default_factory=dict)

    @property
    def latency(self) -> float:
        return self.finished - self.arrival


def _host_set(h) -> dict:
    return {"rows": np.ascontiguousarray(h.rows), "off": np.ascontiguousarray(h.row_offsets),
            "begin": np.ascontiguousarray(h.begin), "end": np.ascontiguousarray(h.end)}


@functools.lru_cache(maxsize=256)
def _disc_layout(n_fine: int, layout: int):
    return workloads.stencil_discs_config(n=n_fine, seed=layout)


def advect_request(rid: int, arrival: float, *, n_fine: int, steps: int, seed: int,
                   velocity=(1.0, 0.5), dtype=torch.float32, layout: Optional[int] = None) -> Request:
    """A disc-refined two-level instance: fine n_fine^2 domain covered ~30% by seeded
    discs + annulus, coarse (n_fine/2)^2 full level, ratio 2, initial values
    U[0, 1] (seeded), ``steps`` upwind + transfer steps.  ``layout``: the
    disc layout's seed (default ``seed``); layouts are cached per size."""
    discs, scale, cov = _disc_layout(int(n_fine), int(seed if layout is None else layout))
    nc = n_fine // 2
    fine_cells = 4 * int((discs.end.astype(np.int64) - discs.begin).sum())  # upper bound of prolong(restrict)
    esize = 4 if dtype == torch.float32 else 8
    est = 2 * esize * (nc * nc + int(1.2 * fine_cells)) + 64 * discs.n + (1 << 20)
    return Request(rid, "advect", arrival, {"discs": _host_set(discs), "n_fine": n_fine, "steps": steps,
                                            "seed": seed, "velocity": tuple(velocity), "dtype": dtype}, est)


def query_request(rid: int, arrival: float, *, rows: int, seed: int, width: int = 65536) -> Request:
    """Two seeded sets of the set-algebra distribution (Poisson(32) intervals
    per row, geometric lengths / gaps 40 / 60) on ``rows`` rows."""
    a = workloads.random_rows_set(rows, width, seed=seed)
    b = workloads.random_rows_set(rows, width, seed=seed + 1)
    est = 16 * (a.n + b.n) * 3 + 16 * rows * 4
    return Request(rid, "query", arrival, {"a": _host_set(a), "b": _host_set(b), "rows": rows}, est)


def make_trace(rate: float, count: int, *, seed: int = 5, advect_frac: float = 0.7, n_fine=(256, 8192),
               steps=(10, 200), query_rows=(256, 16384), size_grid: int = 4, layouts: int = 4) -> List[Request]:
    """Poisson arrivals at ``rate`` requests/s; advection jobs with fine
    domains log-uniform in ``n_fine`` and steps uniform in ``steps``;
    queries with rows uniform in ``query_rows``.  Fine sizes are rounded to
    a log grid of ``size_grid`` points per octave (even sizes) and each job
    takes one of ``layouts`` seeded disc layouts of its size (its initial
    values have their own seed), so that the host builds each geometry
    once; ``size_grid = 0`` keeps every size."""
    rng = np.random.default_rng(seed)
    t = 0.0
    out = []
    lo, hi = np.log2(n_fine[0]), np.log2(n_fine[1])
    for rid in range(count):
        t += float(rng.exponential(1.0 / rate))
        if rng.random() < advect_frac:
            e = rng.uniform(lo, hi)
            if size_grid > 0:
                e = min(max(round(e * size_grid) / size_grid, lo), hi)
            nf = max(2, int(round(2.0 ** e)) // 2 * 2)
            jseed = int(rng.integers(1 << 30))
            out.append(advect_request(rid, t, n_fine=nf, steps=int(rng.integers(steps[0], steps[1] + 1)),
                                      seed=jseed, layout=(1 + jseed % layouts) if layouts > 0 else None))
        else:
            out.append(query_request(rid, t, rows=int(rng.integers(query_rows[0], query_rows[1] + 1)),
                                     seed=int(rng.integers(1 << 30))))
    return out


# --------------------------------------------------------------- scheduler
class Scheduler:
    """FIFO admission under a device-byte budget."""

    def __init__(self, budget_bytes: int = HBM_BUDGET_BYTES):
        self.budget = int(budget_bytes)
        self.used = 0
        self.waiting: deque = deque()
        self.peak = 0
        self.blocked = set()  # requests that found the budget full at least once

    def offer(self, req: Request) -> None:
        if req.est_bytes > self.budget:
            raise ValueError(f"request {req.rid} needs {req.est_bytes} bytes, more than the whole budget")
        self.waiting.append(req)

    def admit(self) -> List[Request]:
        out = []
        while self.waiting and self.used + self.waiting[0].est_bytes <= self.budget:
            r = self.waiting.popleft()
            self.used += r.est_bytes
            out.append(r)
        for r in self.waiting:  # every request still waiting found the budget full
            self.blocked.add(r.rid)
        self.peak = max(self.peak, self.used)
        return out

    def release(self, req: Request) -> None:
        self.used -= req.est_bytes


# ------------------------------------------------------------ device state
def _device_set(h: dict, dev, row_shift: int = 0) -> IntervalSet:
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev, non_blocking=False)
    rows = h["rows"] + row_shift if row_shift else h["rows"]
    return IntervalSet._trusted(t(h["begin"]), t(h["end"]), t(h["off"]), t(rows.astype(np.int32)), dense=None)


class AdvectionJob:
    """One advection job as the plain loop runs it: its own CompositeTwoLevel,
    eager steps."""

    def __init__(self, req: Request, dev):
        from .amr2.composite import CompositeTwoLevel
        from .multilevel import restrict_set

        p = req.payload
        self.req = req
        self.remaining = int(p["steps"])
        discs = _device_set(p["discs"], dev)
        nc = p["n_fine"] // 2
        self.comp = CompositeTwoLevel(nc, nc, restrict_set(discs, 2), dtype=p["dtype"], velocity=p["velocity"])
        # initial values U[0, 1) from a seeded device generator (a cohort
        # draws its instances' values the same way)
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(p["seed"]))
        cv = torch.rand(nc * nc, generator=gen, device=dev, dtype=torch.float32)
        fv = torch.rand(self.comp.fine_field.values.numel(), generator=gen, device=dev, dtype=torch.float32)
        self.comp.set_values(cv, fv)
        self.cells_per_step = self.comp.stats().cell_updates_per_step

    def step(self) -> None:
        self.comp.step()
        self.remaining -= 1

    def output(self) -> dict:
        return {"mass": self.comp.mass(), "coarse": self.comp.coarse.values, "fine": self.comp.fine_field.values}


class Cohort:
    """Advection jobs advanced together as one StackedComposite on one
    stream; results are copied out at each job's own step count."""

    def __init__(self, reqs: List[Request], dev, stream: torch.cuda.Stream, use_graph: bool):
        from .amr2.stacked import StackedComposite

        self.reqs = reqs
        self.stream = stream
        self.use_graph = use_graph
        self.event: Optional[torch.cuda.Event] = None
        self.step = 0
        self.steps = [int(r.payload["steps"]) for r in reqs]
        self.done_steps = sorted(set(self.steps))
        self.snaps: Dict[int, dict] = {}
        self.collected = set()
        with torch.cuda.stream(stream):
            p0 = reqs[0].payload
            self.comp = StackedComposite([{"discs": r.payload["discs"], "n_fine": r.payload["n_fine"],
                                           "velocity": r.payload["velocity"]} for r in reqs],
                                         dtype=p0["dtype"], device=dev)
            # initial values U[0, 1) per instance from its seeded generator,
            # drawn exactly as a job alone draws them
            for j, r in enumerate(reqs):
                gen = torch.Generator(device=dev)
                gen.manual_seed(int(r.payload["seed"]))
                nc = r.payload["n_fine"] // 2
                cv = torch.rand(nc * nc, generator=gen, device=dev, dtype=torch.float32)
                fv = torch.rand(self.comp.ext_cells(j), generator=gen, device=dev, dtype=torch.float32)
                self.comp.set_instance_values(j, cv, fv)
            self.comp.finish_init()

    @property
    def finished(self) -> bool:
        return len(self.collected) == len(self.reqs)

    def launch(self, chunk: int) -> None:
        """Up to ``chunk`` steps, stopping at the next requested step count,
        then copies of the jobs that end there (no host synchronisation)."""
        def snap():
            with torch.cuda.stream(self.stream):
                for j, k in enumerate(self.steps):
                    if k == self.step and j not in self.snaps:
                        self.snaps[j] = self.comp.snapshot(j)
                        self.comp.retire(j)  # no more work for a finished instance

        snap()
        pending = [k for k in self.done_steps if k > self.step]
        if pending:
            n = min(chunk, pending[0] - self.step)
            self.comp.advance(n, self.stream, self.use_graph)
            self.step += n
            snap()
        with torch.cuda.stream(self.stream):
            self.event = torch.cuda.Event()
            self.event.record(self.stream)

    def ready(self) -> bool:
        return self.event is None or self.event.query()

    def collect(self, keep_outputs: bool) -> List[tuple]:
        """(job index, output) of the jobs whose copies are complete (call
        when ``ready``)."""
        out = []
        with torch.cuda.stream(self.stream):
            for j, snap in self.snaps.items():
                if j in self.collected:
                    continue
                o = {"mass": float(snap["mass"].item())}
                if keep_outputs:
                    o["coarse"] = snap["coarse"].cpu().numpy()
                    o["fine"] = snap["fine"].cpu().numpy()
                self.collected.add(j)
                out.append((j, o))
        return out

    def cell_updates(self, j: int) -> int:
        return self.comp.cells_per_step(j) * self.steps[j]


def _query_intervals(r: Request) -> int:
    return int(r.payload["a"]["off"][-1]) + int(r.payload["b"]["off"][-1])


def _stack_queries(reqs: List[Request], dev):
    """Row-stack the A and B sets of several queries (16 empty rows between
    queries); returns device sets and each query's row range.  The stacked
    arrays are written once into pinned host buffers and copied to the
    device asynchronously on the current stream, so the host goes on at once
    (a copy from pageable memory would hold it for the whole transfer)."""
    gap = 16
    shift = 0
    parts_a, parts_b, ranges = [], [], []
    for r in reqs:
        a, b = r.payload["a"], r.payload["b"]
        parts_a.append((a, shift))
        parts_b.append((b, shift))
        ranges.append((shift, r.payload["rows"]))
        shift += r.payload["rows"] + gap

    def cat(parts):
        nrows = sum(h["rows"].size for h, _ in parts)
        nint = sum(int(h["off"][-1]) for h, _ in parts)
        host = {k: torch.empty(n, dtype=torch.int32, pin_memory=True)
                for k, n in (("rows", nrows), ("off", nrows + 1), ("begin", nint), ("end", nint))}
        hv = {k: t.numpy() for k, t in host.items()}
        hv["off"][0] = 0
        r0, i0 = 0, 0
        for h, s in parts:
            nr, ni = h["rows"].size, int(h["off"][-1])
            np.add(h["rows"], s, out=hv["rows"][r0:r0 + nr], casting="unsafe")
            np.add(h["off"][1:], i0, out=hv["off"][r0 + 1:r0 + nr + 1], casting="unsafe")
            hv["begin"][i0:i0 + ni] = h["begin"]
            hv["end"][i0:i0 + ni] = h["end"]
            r0 += nr
            i0 += ni
        d = {k: t.to(dev, non_blocking=True) for k, t in host.items()}
        return IntervalSet._trusted(d["begin"], d["end"], d["off"], d["rows"], dense=None)

    return cat(parts_a), cat(parts_b), ranges


class QueryBatch:
    """Set-algebra queries answered together on one stream: one fused launch
    per operation over the row-stacked pairs (rows of the pairs are identical
    per query, so no row alignment is needed), outputs copied to pinned host
    buffers without a host synchronisation.  ``collect`` splits them per
    query once ``ready``."""

    def __init__(self, reqs: List[Request], dev, stream: torch.cuda.Stream):
        self.reqs = reqs
        self.stream = stream
        self.host = {}
        self._keep = []
        with torch.cuda.stream(stream):
            A, B, self.ranges = _stack_queries(reqs, dev)
            R = A.row_count
            cap = max(A.interval_count + B.interval_count, 1)  # bounds every operation's output
            for op, name in ((0, "union"), (1, "intersection"), (2, "difference")):
                out_off, ob, oe, ctx = _core.setop_raw(op, (A.begin, A.end, A.row_offsets),
                                                       (B.begin, B.end, B.row_offsets), nrows=R, capacity=cap)
                h = tuple(torch.empty(t.numel(), dtype=torch.int32, pin_memory=True) for t in (out_off, ob, oe))
                for dst, src in zip(h, (out_off, ob, oe)):
                    dst.copy_(src, non_blocking=True)
                self.host[name] = h
                self._keep += [out_off, ob, oe]
            self._keep += [A, B]
            # status word of the three launches, read without a host wait
            self.status = torch.empty(1, dtype=torch.int32, pin_memory=True)
            self.status.copy_(ctx.words[:1], non_blocking=True)
            self.ctx = ctx
            self.event = torch.cuda.Event()
            self.event.record(stream)

    def ready(self) -> bool:
        return self.event.query()

    def collect(self, keep_outputs: bool = True) -> List[dict]:
        """Per-query results (empty dicts unless ``keep_outputs``); raises if
        a launch flagged an error (call when ``ready``)."""
        if int(self.status[0]):
            with torch.cuda.stream(self.stream):
                self.ctx.check_status()  # raises with the flags' meaning
        if not keep_outputs:
            self._keep = []
            return [{} for _ in self.reqs]
        outs = {name: tuple(t.numpy() for t in h) for name, h in self.host.items()}
        res = []
        k = 0
        for (r0, nr) in self.ranges:
            one = {}
            for name, (off, ob, oe) in outs.items():
                o = off[k:k + nr + 1].copy()
                one[name] = (o - o[0], ob[o[0]:o[-1]].copy(), oe[o[0]:o[-1]].copy())
            res.append(one)
            k += nr  # the gap between queries is in the row ids only, not in the CSR rows
        self._keep = []
        return res


def run_queries(reqs: List[Request], dev, keep_outputs: bool = True) -> List[dict]:
    """A batch of queries on the current stream, waiting for the results."""
    if not reqs:
        return []
    qb = QueryBatch(reqs, dev, torch.cuda.current_stream(dev))
    qb.event.synchronize()
    return qb.collect(keep_outputs)


# ------------------------------------------------------------------ server
@dataclass
class ServeStats:
    results: List[Result]
    wall: float
    cell_updates: int
    query_intervals: int
    iterations: int
    peak_bytes: int
    setup_s: float = 0.0
    blocked: int = 0

    def summary(self) -> dict:
        lat = {k: sorted(r.latency for r in self.results if r.kind == k) for k in ("advect", "query")}
        allat = sorted(r.latency for r in self.results)
        queue = sorted(r.admitted - r.arrival for r in self.results)
        service = sorted(r.finished - r.admitted for r in self.results)

        def pct(v, q):
            return float(np.percentile(v, q)) if v else None

        return {
            "queue_delay_p50_s": pct(queue, 50), "queue_delay_p99_s": pct(queue, 99),
            "service_p50_s": pct(service, 50), "service_p99_s": pct(service, 99),
            "budget_blocked_requests": self.blocked,
            "requests": len(self.results), "wall_s": self.wall,
            "throughput_req_per_s": len(self.results) / self.wall,
            "cell_updates_per_s": self.cell_updates / self.wall,
            "query_interval_ops_per_s": self.query_intervals / self.wall,
            "latency_p50_s": pct(allat, 50), "latency_p99_s": pct(allat, 99),
            "advect_p50_s": pct(lat["advect"], 50), "advect_p99_s": pct(lat["advect"], 99),
            "query_p50_s": pct(lat["query"], 50), "query_p99_s": pct(lat["query"], 99),
            "iterations": self.iterations, "peak_admitted_bytes": self.peak_bytes, "job_setup_s": self.setup_s,
        }


class Server:
    """Continuous batching server (see the module docstring).  ``window``:
    seconds an admitted advection job may wait for others to join its cohort
    while other cohorts run; ``max_cohort``: instances per cohort;
    ``max_cohort_bytes``: estimated bytes per cohort (the estimate bounds the
    cohort's fields well under the 4 GiB a level may hold for the segment
    stencil's 32-bit byte offsets); ``query_batch``: input intervals per
    query batch."""

    def __init__(self, *, budget_bytes: int = HBM_BUDGET_BYTES, streams: int = 4, chunk: int = 8,
                 use_graph: bool = True, device=None, keep_outputs: bool = False, window: float = 2e-3,
                 max_cohort: int = 64, max_cohort_bytes: int = 8 << 30, query_batch: int = 1 << 21):
        self.dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
        self.sched = Scheduler(budget_bytes)
        self.streams = [torch.cuda.Stream(self.dev) for _ in range(max(1, streams))]
        self.chunk = max(1, int(chunk))
        self.use_graph = use_graph
        self.keep_outputs = keep_outputs
        self.window = float(window)
        self.max_cohort = max(1, int(max_cohort))
        self.max_cohort_bytes = int(max_cohort_bytes)
        self.query_batch = int(query_batch)
        self.qstream = torch.cuda.Stream(self.dev)

    def run(self, trace: List[Request]) -> ServeStats:
        pending = deque(sorted(trace, key=lambda r: r.arrival))
        cohorts: List[Cohort] = []
        qbatches: List[QueryBatch] = []
        waiting: List[Request] = []  # admitted advection jobs without a cohort yet
        qwait: deque = deque()  # admitted queries not yet in a batch
        wait_since = 0.0
        admitted_at: Dict[int, float] = {}
        results: List[Result] = []
        cell_updates = 0
        qints = 0
        iters = 0
        rr = 0
        setup = 0.0
        main = torch.cuda.current_stream(self.dev)
        torch.cuda.synchronize(self.dev)
        t0 = time.perf_counter()
        now = lambda: time.perf_counter() - t0

        while pending or cohorts or qbatches or waiting or qwait or self.sched.waiting:
            t = now()
            while pending and pending[0].arrival <= t:
                self.sched.offer(pending.popleft())
            progressed = False
            for r in self.sched.admit():
                admitted_at[r.rid] = now()
                if r.kind == "query":
                    qwait.append(r)
                else:
                    if not waiting:
                        wait_since = now()
                    waiting.append(r)
            if waiting and (not cohorts or len(waiting) >= self.max_cohort or now() - wait_since >= self.window):
                # shortest jobs first, so that a cohort's jobs end close together (an instance that has
                # ended still costs its cohort transfers until the whole cohort retires)
                waiting.sort(key=lambda r: (int(r.payload["steps"]), r.arrival))
                by_dtype: Dict[torch.dtype, List[Request]] = {}
                take, nbytes = 0, 0
                while (take < len(waiting) and take < self.max_cohort
                       and (take == 0 or nbytes + waiting[take].est_bytes <= self.max_cohort_bytes)):
                    nbytes += waiting[take].est_bytes
                    take += 1
                for r in waiting[:take]:
                    by_dtype.setdefault(r.payload["dtype"], []).append(r)
                waiting = waiting[take:]
                wait_since = now()
                for reqs in by_dtype.values():
                    s = self.streams[rr % len(self.streams)]
                    rr += 1
                    s.wait_stream(main)
                    ts = time.perf_counter()
                    c = Cohort(reqs, self.dev, s, self.use_graph)
                    setup += time.perf_counter() - ts
                    c.launch(self.chunk)
                    cohorts.append(c)
                progressed = True
            # cohorts whose last chunk finished: results, next chunk, retire
            still = []
            for c in cohorts:
                if not c.ready():
                    still.append(c)
                    continue
                for j, out in c.collect(self.keep_outputs):
                    r = c.reqs[j]
                    cell_updates += c.cell_updates(j)
                    results.append(Result(r.rid, "advect", r.arrival, admitted_at[r.rid], now(), out))
                if c.finished:
                    for r in c.reqs:
                        self.sched.release(r)
                else:
                    iters += 1
                    c.launch(self.chunk)
                    still.append(c)
                progressed = True
            cohorts = still
            # queries after the cohorts, at most query_batch intervals per loop
            # iteration: a batch's host work (stacking, copies) stays short,
            # so the cohorts' next chunks are never held back for long
            if qwait:
                queries, nint = [], 0
                while qwait and (not queries or nint + _query_intervals(qwait[0]) <= self.query_batch):
                    nint += _query_intervals(qwait[0])
                    queries.append(qwait.popleft())
                iters += 1
                self.qstream.wait_stream(main)
                qbatches.append(QueryBatch(queries, self.dev, self.qstream))
                progressed = True
            still_q = []
            for qb in qbatches:
                if not qb.ready():
                    still_q.append(qb)
                    continue
                outs = qb.collect(self.keep_outputs)
                t_q = now()
                for r, o in zip(qb.reqs, outs):
                    qints += 3 * (int(r.payload["a"]["off"][-1]) + int(r.payload["b"]["off"][-1]))
                    results.append(Result(r.rid, "query", r.arrival, admitted_at[r.rid], t_q, o))
                    self.sched.release(r)
                progressed = True
            qbatches = still_q
            if not progressed:
                nxt = pending[0].arrival - now() if pending else 0.0
                time.sleep(min(max(nxt, 0.0), 2e-4) if (cohorts or qbatches or waiting or nxt > 0) else 0.0)
        torch.cuda.synchronize(self.dev)
        wall = now()
        st = ServeStats(sorted(results, key=lambda r: r.rid), wall, cell_updates, qints, iters, self.sched.peak)
        st.setup_s = setup
        st.blocked = len(self.sched.blocked)
        return st


def plain_loop(trace: List[Request], device=None, keep_outputs: bool = False) -> ServeStats:
    """Unbatched plain form: requests one at a time in arrival order, eager
    launches, a host synchronisation per step (advection) and per query."""
    dev = torch.device(device) if device is not None else torch.device("cuda", torch.cuda.current_device())
    results = []
    cell_updates = 0
    qints = 0
    setup = 0.0
    torch.cuda.synchronize(dev)
    t0 = time.perf_counter()
    now = lambda: time.perf_counter() - t0
    for r in sorted(trace, key=lambda r: r.arrival):
        while now() < r.arrival:
            time.sleep(min(r.arrival - now(), 0.01))
        ta = now()
        if r.kind == "query":
            o = run_queries([r], dev, keep_outputs)[0]
            qints += 3 * (int(r.payload["a"]["off"][-1]) + int(r.payload["b"]["off"][-1]))
            results.append(Result(r.rid, "query", r.arrival, ta, now(), o if keep_outputs else {}))
        else:
            ts = time.perf_counter()
            job = AdvectionJob(r, dev)
            setup += time.perf_counter() - ts
            while job.remaining:
                job.step()
                torch.cuda.synchronize(dev)
                cell_updates += job.cells_per_step
            out = job.output()
            out = ({"mass": out["mass"], "coarse": out["coarse"].cpu().numpy(), "fine": out["fine"].cpu().numpy()}
                   if keep_outputs else {"mass": out["mass"]})
            results.append(Result(r.rid, "advect", r.arrival, ta, now(), out))
    return ServeStats(results, now(), cell_updates, qints, len(trace), 0, setup)
