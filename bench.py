"""Benchmark driver: one JSON line on stdout.

Headline (N=1): the set-algebra microbenchmark configuration (two seeded
interval sets of 4096 rows x 65536, Poisson(32) intervals per row, geometric
lengths/gaps with means 40/60), batched: B independent seeded pairs stacked
row-wise and processed by ONE fused kernel launch per operation.  A step =
union + intersection + difference over the batch, inputs resident in HBM, no
host synchronisation inside the step.  value = input intervals processed per
second (the reference's ns-per-input-interval convention, inverted).

Also reported beside the headline (each in its own object):
  copy           1 GiB device copy: the achievable-bandwidth reference and the
                 calibration kernel for the HBM byte counters
  plain_form     one pair through the reference-shaped evaluate() call
  morphology     dilate / erode, square SE radius 1, 2, 4 (one pair)
  large_stencil  8192^2 fine / 4096^2 coarse composite, 500 steps, FP32 vs FP64
  batched_amr    512 three-level instances, 200 steps, regrid every 8, hipGraph
                 steps, vs the plain single-instance form on a sample
  plain_amr      reference-shaped two-level run, 128^2 Gaussian bump, FP64
  serve          served stream (Poisson arrivals, continuous batching under an
                 HBM budget) vs the unbatched plain loop: throughput, p50/p99
  roofline       dominant kernel (fused union) against 8 TB/s
  cpu_baseline   the C oracle (OpenMP over rows) on a bounded sample
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
SECTIONS = ("copy", "morphology", "large_stencil", "batched_amr", "plain_amr", "serve")


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=20)
    p.add_argument("--warmup", type=int, default=3)
    p.add_argument("--batch", type=int, default=128, help="independent set pairs per launch")
    p.add_argument("--cpu-seconds", type=float, default=10.0)
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--group", type=int, default=None, help="override lane-group size")
    p.add_argument("--sections", default=",".join(SECTIONS),
                   help="comma-separated extra sections (" + ",".join(SECTIONS) + ") or 'none'")
    p.add_argument("--stencil-steps", type=int, default=500)
    p.add_argument("--amr-instances", type=int, default=512)
    p.add_argument("--amr-steps", type=int, default=200)
    p.add_argument("--serve-rates", default="100,1000,10000", help="Poisson arrival rates (requests/s) for the serve section")
    p.add_argument("--serve-requests", type=int, default=256)
    p.add_argument("--plain-reps", type=int, default=50, help="repetitions of the plain single-pair form")
    return p.parse_args()


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def measured_traffic(batch):
    """HBM bytes per fused-union launch from the committed rocprofv3 summary
    of this workload (FETCH_SIZE / WRITE_SIZE, copy-kernel calibrated), or
    None when the summary does not cover the configuration run."""
    path = os.path.join(ROOT, "profiles", "setop_union_batched128.md")
    if batch != 128 or not os.path.exists(path):
        return None
    for line in open(path):
        if line.startswith("traffic_bytes_per_launch:"):
            return int(line.split(":")[1])
    return None


def cpu_baseline(pair, seconds):
    from oracle import oracle as ox

    threads = ox.num_threads()
    oa = ox.OSet(pair[0].rows, pair[0].row_offsets, pair[0].begin, pair[0].end)
    ob = ox.OSet(pair[1].rows, pair[1].row_offsets, pair[1].begin, pair[1].end)
    done = 0
    reps = 0
    t0 = time.perf_counter()
    while time.perf_counter() - t0 < seconds:
        for op in ("union", "intersection", "difference"):
            ox.setop(op, oa, ob)
            done += oa.n + ob.n
        reps += 1
    dt = time.perf_counter() - t0
    return {"value": done / dt, "unit": "intervals/s", "cores": threads, "kind": "port",
            "sample": f"one 4096x65536 pair, union+intersection+difference repeated {reps}x ({dt:.1f}s), "
                      "C oracle with OpenMP over rows"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py: this build measures one GPU; --gpus {args.gpus} is not supported "
              "(the interval-set path shards by rows / instances with no communication)", file=sys.stderr)
        sys.exit(2)
    import torch

    from subsetix_amd import perf

    torch.cuda.set_device(0)
    dev = torch.device("cuda", 0)
    sections = [] if args.sections == "none" else [s for s in args.sections.split(",") if s]
    for s in sections:
        if s not in SECTIONS:
            sys.exit(f"bench.py: unknown section {s!r}")

    t0 = time.time()
    head = perf.setalgebra_batched(args.batch, args.steps, args.warmup, dev, group=args.group)
    log(f"set algebra ({time.time() - t0:.1f}s): kernel ms {head['kernel_ms']} GB/s {head['kernel_gbs']}")
    pair = head.pop("host_pair")
    plain = perf.setalgebra_plain(pair, args.plain_reps, dev)
    achieved = head["kernel_gbs"]["union"]

    extra = {}
    for s in sections:
        t0 = time.time()
        if s == "copy":
            extra[s] = perf.copy_bandwidth(dev)
        elif s == "morphology":
            extra[s] = perf.morphology(pair, 10, dev)
        elif s == "large_stencil":
            extra[s] = perf.large_stencil(args.stencil_steps, dev)
        elif s == "batched_amr":
            extra[s] = perf.batched_amr(args.amr_instances, args.amr_steps, dev)
        elif s == "plain_amr":
            extra[s] = perf.plain_two_level(dev)
        elif s == "serve":
            extra[s] = perf.serve([float(r) for r in args.serve_rates.split(",")], args.serve_requests, dev)
        torch.cuda.empty_cache()
        log(f"{s} ({time.time() - t0:.1f}s): {json.dumps(extra[s])}")

    cpu = None if args.no_cpu else cpu_baseline(pair, args.cpu_seconds)

    out = {
        "metric": "interval-ops/s",
        "value": head["interval_ops_per_s"],
        "unit": "intervals/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": head["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "int32",
        "data": "synthetic (seeded)",
        "config": {
            "workload": "set-algebra microbenchmark, batched: union+intersection+difference over "
                        f"{args.batch} stacked seeded pairs of 4096 rows x 65536 (Poisson(32) intervals/row)",
            "pairs": args.batch,
            "rows": head["rows"],
            "input_intervals_per_op": head["input_intervals_per_op"],
            "output_intervals": head["outputs"],
        },
        "kernel_ms": head["kernel_ms"],
        "kernel_gbs": head["kernel_gbs"],
        "plain_form": {"value": plain["interval_ops_per_s"], "unit": "intervals/s", "desc": plain["desc"],
                       "ms_per_op": plain["ms_per_op"]},
        "batched_vs_plain": head["interval_ops_per_s"] / plain["interval_ops_per_s"],
        "autotune": head.get("autotune"),
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": measured_traffic(args.batch),
                     "kernel": "setop fused union",
                     "algorithmic_bytes": head["algorithmic_bytes"]["union"]},
        "cpu_baseline": cpu,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
