"""Benchmark driver: one JSON line on stdout.

Default workload (N=1): the set-algebra microbenchmark configuration (two
seeded interval sets of 4096 rows x 65536, Poisson(32) intervals per row,
geometric lengths/gaps with means 40/60), batched: B independent seeded
pairs stacked row-wise and processed by ONE fused kernel launch per
operation.  A step = union + intersection + difference over the batch,
inputs resident in HBM, no host synchronisation inside the step.

value = input intervals processed per second (the reference's
ns-per-input-interval convention, inverted).  Also reported: the plain
single-pair form (one pair, reference-shaped call with one host read per
operation), the roofline of the dominant kernel (fused union), and the CPU
baseline (the C oracle, OpenMP over rows, on a bounded sample).
"""

from __future__ import annotations

import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--batch", type=int, default=128, help="independent set pairs per launch")
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--group", type=int, default=None, help="override lane-group size")
    return p.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def setop_bytes(na, nb, nout, ra, rb, rout):
    return 8 * (na + nb + nout) + 4 * (ra + rb + rout + 3)


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the interval-set path shards by rows / instances with no communication; see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)
    import torch

    from subsetix_amd import _core, workloads
    from subsetix_amd.expressions import IntervalSet

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)

    # ---- inputs -----------------------------------------------------------
    t0 = time.time()
    pairs = [workloads.setalgebra_pair(seed_a=1 + 2 * k, seed_b=2 + 2 * k) for k in range(args.batch)]
    A = workloads.stack_rows([p[0] for p in pairs])
    B = workloads.stack_rows([p[1] for p in pairs])
    log(f"generated {args.batch} pairs in {time.time() - t0:.1f}s: NA={A.n} NB={B.n} rows={A.rows.size}")

    def dev_set(h):
        t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
        return IntervalSet._trusted(t(h.begin), t(h.end), t(h.row_offsets), t(h.rows), dense=True)

    a, b = dev_set(A), dev_set(B)
    R = a.row_count
    na, nb = a.interval_count, b.interval_count
    cap = na + nb
    bufs = {op: (torch.empty(cap, dtype=torch.int32, device=dev), torch.empty(cap, dtype=torch.int32, device=dev))
            for op in range(3)}
    totals = {}

    def run_op(op):
        out_off, ob, oe, ctx = _core.setop_raw(op, (a.begin, a.end, a.row_offsets), (b.begin, b.end, b.row_offsets),
                                               nrows=R, capacity=cap, buffers=bufs[op], group=args.group)
        return ctx

    # first pass: output sizes + status check
    for op in range(3):
        ctx = run_op(op)
        (totals[op],) = _core.read_counts(ctx, ctx.words[1])
    log(f"outputs: union={totals[0]} inter={totals[1]} diff={totals[2]}")

    for _ in range(args.warmup):
        for op in range(3):
            run_op(op)
    torch.cuda.synchronize()

    # ---- timed region -----------------------------------------------------
    ev = {op: [] for op in range(3)}
    torch.cuda.synchronize()
    t_start = time.perf_counter()
    for _ in range(args.steps):
        for op in range(3):
            s = torch.cuda.Event(enable_timing=True)
            e = torch.cuda.Event(enable_timing=True)
            s.record()
            run_op(op)
            e.record()
            ev[op].append((s, e))
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_start
    ctx = run_op(0)
    _core.read_counts(ctx)  # raises if any launch flagged an error

    kernel_ms = {op: float(np.mean([s.elapsed_time(e) for s, e in ev[op]])) for op in range(3)}
    ops_per_step = 3 * (na + nb)
    ms_per_step = wall * 1e3 / args.steps
    value = ops_per_step * args.steps / wall
    bytes_union = setop_bytes(na, nb, totals[0], R, R, R)
    achieved = bytes_union / (kernel_ms[0] * 1e-3) / 1e9
    log(f"kernel ms: {kernel_ms}  wall ms/step {ms_per_step:.3f}")

    # ---- plain single-pair form (reference-shaped: one host read per op) --
    import subsetix_amd as sx

    pa = sx.build_interval_set(pairs[0][0].row_offsets, pairs[0][0].begin, pairs[0][0].end)
    pb = sx.build_interval_set(pairs[0][1].row_offsets, pairs[0][1].begin, pairs[0][1].end)
    mk = [sx.make_union, sx.make_intersection, sx.make_difference]
    for _ in range(3):
        for f in mk:
            sx.evaluate(f(sx.make_input(pa), sx.make_input(pb)))
    torch.cuda.synchronize()
    reps = 50
    tp = time.perf_counter()
    for _ in range(reps):
        for f in mk:
            sx.evaluate(f(sx.make_input(pa), sx.make_input(pb)))
    torch.cuda.synchronize()
    plain_s = (time.perf_counter() - tp) / reps
    plain_value = 3 * (pa.interval_count + pb.interval_count) / plain_s

    # ---- CPU baseline (oracle, OpenMP over rows) ---------------------------
    cpu = None
    if not args.no_cpu:
        from oracle import oracle as ox

        threads = ox.num_threads()
        oa = ox.OSet(pairs[0][0].rows, pairs[0][0].row_offsets, pairs[0][0].begin, pairs[0][0].end)
        obs = ox.OSet(pairs[0][1].rows, pairs[0][1].row_offsets, pairs[0][1].begin, pairs[0][1].end)
        done = 0
        tc = time.perf_counter()
        reps_cpu = 0
        while time.perf_counter() - tc < args.cpu_seconds:
            for op in ("union", "intersection", "difference"):
                ox.setop(op, oa, obs)
                done += oa.n + obs.n
            reps_cpu += 1
        cpu_s = time.perf_counter() - tc
        cpu = {"value": done / cpu_s, "unit": "intervals/s", "cores": threads, "kind": "port",
               "sample": f"one 4096x65536 pair, union+intersection+difference repeated {reps_cpu}x "
                         f"({cpu_s:.1f}s), C oracle with OpenMP over rows"}

    out = {
        "metric": "interval-ops/s",
        "value": value,
        "unit": "intervals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded)",
        "config": {
            "workload": "set-algebra microbenchmark, batched: union+intersection+difference over "
                        f"{args.batch} stacked seeded pairs of 4096 rows x 65536 (Poisson(32) intervals/row)",
            "pairs": args.batch,
            "rows": R,
            "input_intervals_per_op": na + nb,
            "output_intervals": {"union": totals[0], "intersection": totals[1], "difference": totals[2]},
        },
        "kernel_ms": {"union": kernel_ms[0], "intersection": kernel_ms[1], "difference": kernel_ms[2]},
        "plain_form": {"value": plain_value, "unit": "intervals/s",
                       "desc": "one pair, reference-shaped evaluate() with one host read per op",
                       "ms_per_op": plain_s * 1e3 / 3},
        "batched_vs_plain": value / plain_value,
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None, "kernel": "setop fused union",
                     "algorithmic_bytes": bytes_union},
        "cpu_baseline": cpu,
    }
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
