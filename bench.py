"""Benchmark driver: one JSON line on stdout.

Headline (N=1): the set-algebra microbenchmark configuration (two seeded
interval sets of 4096 rows x 65536, Poisson(32) intervals per row, geometric
lengths/gaps with means 40/60), batched: B independent seeded pairs stacked
row-wise and processed by ONE fused kernel launch per operation.  A step =
union + intersection + difference over the batch, inputs resident in HBM, no
host synchronisation inside the step.  value = input intervals processed per
second (the reference's ns-per-input-interval convention, inverted), from the
wall time of the K timed steps; a second pass launches each operation K times
back to back between two HIP events for the per-kernel times (kernel_ms, the
roofline).

Also reported beside the headline (each in its own object):
  copy           1 GiB device copy: the achievable-bandwidth reference and the
                 calibration kernel for the HBM byte counters
  plain_form     one pair through the reference-shaped evaluate() call
  morphology     dilate / erode, square SE radius 1, 2, 4 (one pair), and dilation /
                 erosion of 64 instances stacked with 16-row gaps in one call
                 (and of 32 instances alternating 16 / 64 intervals per row)
  large_stencil  8192^2 fine / 4096^2 coarse composite, 500 steps, FP32 vs FP64
  batched_amr    512 three-level instances, 200 steps, regrid every 8, device
                 regrid, each regrid + 8 steps one hipGraph replay; exclusive
                 regrid / step GPU times; vs the plain single-instance form
  plain_amr      reference-shaped two-level run, 128^2 Gaussian bump, FP64
  serve          served stream (Poisson arrivals, continuous batching under an
                 HBM budget) vs the unbatched plain loop: throughput, p50/p99
  roofline       dominant kernel (fused union) against 8 TB/s
  cpu_baseline   the C oracle (OpenMP over rows) on a bounded sample, all
                 threads and 1 thread, CPU model recorded
  cpu_baselines  per section: the oracle's restatement of the same workload
                 on the host (bounded sample, all threads and 1 thread) and
                 the GPU FP32 result against the FP64 oracle (max / L2 /
                 mass)
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


def cpu_info():
    """Host CPU model and core counts (the CPU baselines' hardware)."""
    model = None
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    model = line.split(":", 1)[1].strip()
                    break
    except OSError:
        pass
    try:
        usable = len(os.sched_getaffinity(0))
    except AttributeError:
        usable = os.cpu_count()
    return {"model": model, "nproc": os.cpu_count(), "usable": usable}


def _with_threads(ox, n, fn):
    keep = ox.num_threads()
    ox.set_threads(n)
    try:
        t0 = time.perf_counter()
        out = fn()
        return out, time.perf_counter() - t0
    finally:
        ox.set_threads(keep)


def _err(gpu, ref):
    d = np.asarray(gpu, np.float64) - np.asarray(ref, np.float64)
    return {"max_abs": float(np.abs(d).max()) if d.size else 0.0, "l2_rms": float(np.sqrt(np.mean(d * d))) if d.size
            else 0.0}


def cpu_baseline(pair, seconds):
    """Headline CPU leg: the C oracle's set algebra (the reference's K1/K2
    loops restated, OpenMP over rows) on one pair, all threads and 1 thread."""
    from oracle import oracle as ox

    threads = ox.num_threads()
    oa = ox.OSet(pair[0].rows, pair[0].row_offsets, pair[0].begin, pair[0].end)
    ob = ox.OSet(pair[1].rows, pair[1].row_offsets, pair[1].begin, pair[1].end)

    def run(limit):
        done, reps = 0, 0
        t0 = time.perf_counter()
        while time.perf_counter() - t0 < limit:
            for op in ("union", "intersection", "difference"):
                ox.setop(op, oa, ob)
                done += oa.n + ob.n
            reps += 1
        return done, reps, time.perf_counter() - t0

    done, reps, dt = run(seconds)
    (d1, r1, t1), _ = _with_threads(ox, 1, lambda: run(max(1.0, seconds / 4)))
    return {"value": done / dt, "unit": "intervals/s", "cores": threads, "kind": "port",
            "value_1core": d1 / t1, "cpu": cpu_info(),
            "sample": f"one 4096x65536 pair, union+intersection+difference repeated {reps}x ({dt:.1f}s) on "
                      f"{threads} threads and {r1}x ({t1:.1f}s) on 1 thread; C oracle (OpenMP over rows)"}


def cpu_large_stencil(dev, steps=4):
    """Large stencil CPU leg: the oracle composite step (prolong ghosts,
    upwind both levels, restrict) at the 8192^2 disc geometry, FP32 (the
    reference's precision) timed on all threads and on 1 thread; the GPU
    FP32 run against the FP64 oracle after the same steps."""
    import torch

    from oracle import oracle as ox
    from subsetix_amd import workloads
    from subsetix_amd.amr2.composite import CompositeTwoLevel
    from subsetix_amd.multilevel import restrict_set
    from subsetix_amd.perf import _dev_set

    discs, _, _ = workloads.stencil_discs_config(n=8192, seed=3)
    refine_h = ox.restrict_set(ox.OSet(discs.rows, discs.row_offsets, discs.begin, discs.end), 2)
    W = 4096
    rng = np.random.default_rng(3)
    cv0 = rng.random(W * W).astype(np.float32)
    o32 = ox.Composite(W, W, refine_h, velocity=(1.0, 0.5), cfl=0.5, dtype=np.float32)
    fv0 = rng.random(o32.fv.size).astype(np.float32)
    cells = int(o32.cv.size + o32.fv.size)
    o32.set_values(cv0, fv0)
    _, t_all = _with_threads(ox, ox.num_threads(), lambda: [o32.step() for _ in range(steps)])
    _, t_one = _with_threads(ox, 1, lambda: o32.step())
    o64 = ox.Composite(W, W, refine_h, velocity=(1.0, 0.5), cfl=0.5, dtype=np.float64)
    o64.set_values(cv0.astype(np.float64), fv0.astype(np.float64))
    for _ in range(steps):
        o64.step()
    g = CompositeTwoLevel(W, W, restrict_set(_dev_set(discs, dev), 2), dtype=torch.float32, velocity=(1.0, 0.5),
                          cfl=0.5)
    g.set_values(torch.from_numpy(cv0).to(dev), torch.from_numpy(fv0).to(dev))
    for _ in range(steps):
        g.step()
    m32, m64 = g.mass(), o64.mass()
    return {"value": cells * steps / t_all, "value_1core": cells / t_one, "unit": "cell-updates/s",
            "cores": ox.num_threads(), "kind": "port", "cells_per_step": cells,
            "sample": f"{steps} FP32 composite steps on all threads, 1 step on 1 thread",
            "f32_vs_fp64_oracle": {"steps": steps, "coarse": _err(g.coarse.values.cpu().numpy(), o64.cv),
                                   "fine": _err(g.fine_field.values.cpu().numpy(), o64.fv),
                                   "mass_rel_err": abs(m32 - m64) / abs(m64)}}


def cpu_batched_amr(dev, steps=10):
    """Batched AMR CPU leg: the oracle's three-level instance engine on a
    sample of the 512-instance draw (smallest, median and largest coarse
    size), timed over instances on all threads and 1 thread; the GPU FP32
    batch of the same instances against the FP64 oracle."""
    import torch

    from oracle import oracle as ox
    from subsetix_amd import workloads
    from subsetix_amd.amr3 import DeviceBatchedAMR
    from subsetix_amd.batch import BatchConfig

    insts = workloads.batched_instances(count=512, seed=4)
    ns = np.array([i["n"] for i in insts])
    picks = [int(np.argmin(ns)), int(np.argsort(ns)[ns.size // 2]), int(np.argsort(ns)[int(0.9 * ns.size)]),
             int(np.argmax(ns))]
    sub = [insts[i] for i in picks]

    def oracle_run(dtype):
        outs, cells = [], 0
        for inst in sub:
            o = ox.Instance3(inst["n"], workloads.bumps_field(inst["n"], inst["bumps"], dtype=dtype).reshape(-1),
                             inst["velocity"], dtype=dtype)
            for _ in range(steps):
                o.run(1)
                cells += sum(int(ox.cell_offsets(s)[-1]) for s in o.sets)
            outs.append(o)
        return outs, cells

    (o32, cells), t_all = _with_threads(ox, ox.num_threads(), lambda: oracle_run(np.float32))
    (_, c1), t_one = _with_threads(ox, 1, lambda: oracle_run(np.float32))
    o64, _ = oracle_run(np.float64)
    eng = DeviceBatchedAMR(sub, dtype=torch.float32, config=BatchConfig(levels=3, regrid_every=8,
                                                                          threshold=(0.02, 0.02), grading=1),
                           device=dev)
    eng.run(steps)
    m = eng.masses().cpu().numpy()
    errs = []
    for j, o in enumerate(o64):
        e = _err(eng.level_values(0, j).reshape(-1), o.vals[0])
        e.update({"n": sub[j]["n"], "mass_rel_err": float(abs(m[j] - o.mass()) / abs(o.mass()))})
        errs.append(e)
    return {"value": cells / t_all, "value_1core": c1 / t_one, "unit": "cell-updates/s", "cores": ox.num_threads(),
            "kind": "port", "sample": f"instances {picks} (coarse {[s['n'] for s in sub]}), {steps} steps each "
                                      "(regrid before step 0 and 8), all threads and 1 thread",
            "f32_vs_fp64_oracle_level0": errs}


def cpu_plain_amr(dev):
    """Plain two-level CPU leg: the oracle loop (128^2, 100 steps, regrid
    10) in FP64 on all threads and 1 thread; the GPU FP32 run against it."""
    import torch

    import subsetix_amd as sx
    from oracle import oracle as ox
    from subsetix_amd import workloads
    from subsetix_amd.amr2 import AMR2Simulation, SimulationConfig

    n, steps = 128, 100
    u0 = workloads.gaussian_bump(n).reshape(-1)
    (cv, fv, _), t_all = _with_threads(ox, ox.num_threads(), lambda: ox.two_level_loop(u0, n, steps, 10))
    _, t_one = _with_threads(ox, 1, lambda: ox.two_level_loop(u0, n, steps, 10))
    sim = AMR2Simulation(SimulationConfig(coarse_resolution=n, velocity=(1.0, 0.5), cfl=0.5, refine_threshold=0.1,
                                          grading=2, grading_mode="moore", tag="value", accumulate="f64"))
    f = sx.create_interval_field(sx.full_interval_set(n, n, device=dev), 0.0, dtype=torch.float32)
    f.values.copy_(torch.from_numpy(u0.astype(np.float32)).to(dev))
    sim.set_initial_field(f)
    sim.run(steps=steps, regrid_every=10)
    m32, m64 = sim.total_mass(), ox.mass(cv, 1.0 / n / n)
    cells = 5 * n * n * steps
    return {"value": cells / t_all, "value_1core": cells / t_one, "unit": "cell-updates/s",
            "cores": ox.num_threads(), "kind": "port", "sample": "the whole 100-step FP64 run",
            "f32_vs_fp64_oracle": {"coarse": _err(sim.state.coarse_field.values.cpu().numpy(), cv),
                                   "fine": _err(sim.state.fine_field.values.cpu().numpy(), fv),
                                   "mass_rel_err": abs(m32 - m64) / abs(m64)}}


def cpu_served_jobs(dev, count=3):
    """Served-stream CPU leg: advection jobs of the served trace (fine
    256..8192, seeded) run by the oracle composite in FP64 and FP32 (timed);
    the GPU FP32 job (alone == in the mix, bit-exact) against FP64."""
    import torch

    from oracle import oracle as ox
    from subsetix_amd.serve import AdvectionJob, make_trace

    trace = [r for r in make_trace(1000.0, 64, seed=5) if r.kind == "advect"][:count]
    out, cells, t = [], 0, 0.0
    for r in trace:
        p = r.payload
        nc = p["n_fine"] // 2
        d = p["discs"]
        job = AdvectionJob(r, dev)
        refine = ox.restrict_set(ox.make(d["off"], d["begin"], d["end"], d["rows"]), 2)
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(p["seed"]))
        cv = torch.rand(nc * nc, generator=gen, device=dev, dtype=torch.float32).cpu().numpy()
        fv = None
        res = {}
        for dt in (np.float64, np.float32):
            o = ox.Composite(nc, nc, refine, velocity=p["velocity"], cfl=0.5, dtype=dt)
            if fv is None:
                fv = torch.rand(o.fv.size, generator=gen, device=dev, dtype=torch.float32).cpu().numpy()
            o.set_values(cv.astype(dt), fv.astype(dt))
            t0 = time.perf_counter()
            for _ in range(p["steps"]):
                o.step()
            if dt == np.float32:
                t += time.perf_counter() - t0
                cells += (o.cv.size + o.fv.size) * p["steps"]
            res[dt] = o
        while job.remaining:
            job.step()
        o64 = res[np.float64]
        m32 = job.comp.mass()
        out.append({"n_fine": p["n_fine"], "steps": p["steps"],
                    "coarse": _err(job.comp.coarse.values.cpu().numpy(), o64.cv),
                    "fine": _err(job.comp.fine_field.values.cpu().numpy(), o64.fv),
                    "mass_rel_err": abs(m32 - o64.mass()) / abs(o64.mass())})
    return {"value": cells / t, "unit": "cell-updates/s", "cores": ox.num_threads(), "kind": "port",
            "sample": f"{count} advection jobs of the served trace, FP32, all threads",
            "f32_vs_fp64_oracle": out}


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
            extra[s]["batched"] = perf.morphology_batched(64, 5, dev)
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
    cpu_legs = {}
    if not args.no_cpu:
        legs = {"large_stencil": cpu_large_stencil, "batched_amr": cpu_batched_amr, "plain_amr": cpu_plain_amr,
                "serve": cpu_served_jobs}
        for name, fn in legs.items():
            if name in sections:
                t0 = time.time()
                cpu_legs[name] = fn(dev)
                torch.cuda.empty_cache()
                log(f"cpu leg {name} ({time.time() - t0:.1f}s): {json.dumps(cpu_legs[name])}")

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
        "cpu_baselines": cpu_legs or None,
    }
    out.update(extra)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
