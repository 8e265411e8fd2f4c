"""Workload runners for the benchmark configurations (timing on the HIP
device with events; every runner returns a plain dict).

  * setalgebra_batched  - stacked pairs of 4096 x 65536 sets, fused set ops
  * setalgebra_plain    - one pair through the reference-shaped API
  * morphology          - dilate / erode (square SE r = 1, 2, 4), one pair
  * large_stencil       - 8192^2 fine / 4096^2 coarse composite, FP32 vs FP64
  * batched_amr         - many 3-level instances, device regrid, hipGraph cycles
  * plain_two_level     - the reference-shaped two-level driver (128^2 bump)
"""

from __future__ import annotations

import sys
import time
import warnings
from typing import Dict, List

import numpy as np
import torch

from . import _core, autotune, workloads
from .expressions import IntervalSet

HBM_PEAK_GBS = 8000.0


def log(msg: str) -> None:
    print(f"[perf] {msg}", file=sys.stderr, flush=True)


def _dev_set(h, dev, dense=True) -> IntervalSet:
    t = lambda a: torch.from_numpy(np.ascontiguousarray(a)).to(dev)
    return IntervalSet._trusted(t(h.begin), t(h.end), t(h.row_offsets), t(h.rows), dense=dense)


def _events():
    return torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)


def setop_bytes(na, nb, nout, ra, rb, rout) -> int:
    return 8 * (na + nb + nout) + 4 * (ra + rb + rout + 3)


def setalgebra_batched(batch: int, steps: int, warmup: int, dev, group=None) -> Dict:
    pairs = [workloads.setalgebra_pair(seed_a=1 + 2 * k, seed_b=2 + 2 * k) for k in range(batch)]
    A = workloads.stack_rows([p[0] for p in pairs])
    B = workloads.stack_rows([p[1] for p in pairs])
    a, b = _dev_set(A, dev), _dev_set(B, dev)
    R = a.row_count
    na, nb = a.interval_count, b.interval_count
    cap = na + nb
    bufs = {op: (torch.empty(cap, dtype=torch.int32, device=dev), torch.empty(cap, dtype=torch.int32, device=dev))
            for op in range(3)}

    def run_op(op):
        return _core.setop_raw(op, (a.begin, a.end, a.row_offsets), (b.begin, b.end, b.row_offsets), nrows=R,
                               capacity=cap, buffers=bufs[op], group=group)[3]

    totals = {}
    with autotune.tuning(group is None):
        for op in range(3):
            ctx = run_op(op)
            (totals[op],) = _core.read_counts(ctx, ctx.words[1])
    for _ in range(warmup):
        for op in range(3):
            run_op(op)
    torch.cuda.synchronize()
    # throughput: the steps alone (a timing event costs ~8 us of GPU time,
    # about 10 % of a step with events around every launch)
    t0 = time.perf_counter()
    for _ in range(steps):
        for op in range(3):
            run_op(op)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    # kernel times: per operation, `steps` launches back to back between two
    # HIP events on the launch stream (the average includes the launch gaps)
    kms = {}
    for op in range(3):
        s, e = _events()
        s.record()
        for _ in range(steps):
            run_op(op)
        e.record()
        torch.cuda.synchronize()
        kms[op] = s.elapsed_time(e) / steps
    _core.read_counts(run_op(0))
    byts = {op: setop_bytes(na, nb, totals[op], R, R, R) for op in range(3)}
    return {
        "pairs": batch, "rows": R, "input_intervals_per_op": na + nb,
        "outputs": {"union": totals[0], "intersection": totals[1], "difference": totals[2]},
        "ms_per_step": wall * 1e3 / steps,
        "interval_ops_per_s": 3 * (na + nb) * steps / wall,
        "kernel_ms": {"union": kms[0], "intersection": kms[1], "difference": kms[2]},
        "kernel_gbs": {k: byts[op] / (kms[op] * 1e-3) / 1e9 for k, op in (("union", 0), ("intersection", 1),
                                                                           ("difference", 2))},
        "algorithmic_bytes": {"union": byts[0], "intersection": byts[1], "difference": byts[2]},
        "autotune": autotune_report(("setop",)) if group is None else None,
        "host_pair": pairs[0],
    }


def autotune_report(families) -> Dict:
    """The autotuner's table for some families: choice and candidate times
    (microseconds) per shape bucket."""
    out = {}
    for key, e in autotune.default().table.items():
        if key[0] in families:
            out["/".join(str(k) for k in key)] = {"choice": str(e.choice),
                                                  "us": {k: round(v, 1) for k, v in e.times_us.items()}}
    return out


def setalgebra_plain(pair, reps: int, dev) -> Dict:
    import subsetix_amd as sx

    pa = sx.build_interval_set(pair[0].row_offsets, pair[0].begin, pair[0].end, device=dev)
    pb = sx.build_interval_set(pair[1].row_offsets, pair[1].begin, pair[1].end, device=dev)
    mk = [sx.make_union, sx.make_intersection, sx.make_difference]
    for _ in range(3):
        for f in mk:
            sx.evaluate(f(sx.make_input(pa), sx.make_input(pb)))
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(reps):
        for f in mk:
            sx.evaluate(f(sx.make_input(pa), sx.make_input(pb)))
    torch.cuda.synchronize()
    per = (time.perf_counter() - t0) / reps
    return {"interval_ops_per_s": 3 * (pa.interval_count + pb.interval_count) / per, "ms_per_op": per * 1e3 / 3,
            "desc": "one pair, reference-shaped evaluate() with one host read per operation"}


def morphology(pair, reps: int, dev) -> Dict:
    import subsetix_amd as sx

    s = sx.build_interval_set(pair[0].row_offsets, pair[0].begin, pair[0].end, device=dev)
    out = {}
    for r in (1, 2, 4):
        for name, fn in (("dilate", lambda: sx.dilate_interval_set(s, r, r)),
                         ("erode", lambda: sx.erode_interval_set(s, r, r, width=65536, height=4096))):
            res = fn()
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                res = fn()
            torch.cuda.synchronize()
            per = (time.perf_counter() - t0) / reps
            out[f"{name}_r{r}"] = {"ms": per * 1e3, "out_intervals": res.interval_count,
                                   "interval_ops_per_s": s.interval_count / per}
    return out


def morphology_batched(batch: int, reps: int, dev) -> Dict:
    """Dilation and erosion by the square structuring element of radius 1, 2, 4 on
    ``batch`` set-algebra instances (4096 x 65536, the A sets of the batched
    set-algebra workload) stacked row-wise with 16 empty rows between them,
    one dilation call for the whole stack (dense offsets, the tiled LDS
    window union, row compaction: one host read), and one erosion call over
    the stack's bounding domain 65536 x (last row + 1) (one launch of row
    tiles, one host read; the empty rows between instances are inside that
    domain, so each instance also erodes from its top and bottom rows).  Algorithmic bytes: 8 B per input and output interval and
    4 B per input and output row id and offset.  The same calls on an uneven
    stack ("uneven": ``batch`` / 2 instances of 4096 x 65536 alternating
    Poisson(16) and Poisson(64) intervals per row), where the row tiles'
    count balancing matters.  Two untimed calls before each timed batch (the
    first ones on a shape allocate device memory)."""
    import subsetix_amd as sx

    def timed(s, height):
        out = {"rows": s.row_count, "input_intervals": s.interval_count, "erode_domain": [65536, height]}
        for r in (1, 2, 4):
            for name, fn in (("dilate", lambda: sx.dilate_interval_set(s, r, r)),
                             ("erode", lambda: sx.erode_interval_set(s, r, r, width=65536, height=height))):
                res = fn()
                res = fn()
                torch.cuda.synchronize()
                t0 = time.perf_counter()
                for _ in range(reps):
                    res = fn()
                torch.cuda.synchronize()
                per = (time.perf_counter() - t0) / reps
                byts = 8 * (s.interval_count + res.interval_count) + 8 * (s.row_count + res.row_count)
                out[f"{name}_r{r}"] = {"ms": per * 1e3, "out_intervals": res.interval_count,
                                       "out_rows": res.row_count, "interval_ops_per_s": s.interval_count / per,
                                       "algorithmic_bytes": byts, "gbs": byts / per / 1e9}
        return out

    sets = [workloads.setalgebra_pair(seed_a=1 + 2 * k, seed_b=2 + 2 * k)[0] for k in range(batch)]
    h = workloads.stack_rows_gapped(sets, 16)
    out = {"instances": batch, **timed(_dev_set(h, dev, dense=False), int(h.rows[-1]) + 1)}
    sets = [workloads.random_rows_set(4096, 65536, seed=200 + k, mean_count=64 if k % 2 else 16)
            for k in range(max(batch // 2, 2))]
    h = workloads.stack_rows_gapped(sets, 16)
    out["uneven"] = {"instances": len(sets), **timed(_dev_set(h, dev, dense=False), int(h.rows[-1]) + 1)}
    return out


def large_stencil(steps: int, dev, n: int = 8192, seed: int = 3) -> Dict:
    """Fine n^2 covered ~30% by seeded discs + annulus; coarse (n/2)^2 full.
    FP32 and FP64 runs from the same initial values (uniform [0, 1])."""
    from .amr2.composite import CompositeTwoLevel
    from .multilevel import restrict_set

    t0 = time.perf_counter()
    discs, scale, cov = workloads.stencil_discs_config(n=n, seed=seed)
    gen_s = time.perf_counter() - t0
    log(f"large_stencil: discs generated in {gen_s:.1f}s (coverage {cov:.4f})")
    discs_d = _dev_set(discs, dev)
    refine = restrict_set(discs_d, 2)
    W = H = n // 2
    res = {"fine_domain": n, "coverage": cov, "radius_scale": scale, "generate_s": gen_s}
    runs = {}
    rng = np.random.default_rng(seed)
    cv0 = None
    for name, dt in (("f32", torch.float32), ("f64", torch.float64)):
        comp = CompositeTwoLevel(W, H, refine, dtype=dt, velocity=(1.0, 0.5), cfl=0.5)
        if cv0 is None:
            cv0 = torch.from_numpy(rng.random(W * H).astype(np.float32)).to(dev)
            fv0 = torch.from_numpy(rng.random(comp.fine_field.values.numel()).astype(np.float32)).to(dev)
        comp.set_values(cv0, fv0)
        m0 = comp.mass()
        comp.prepare(tune=True)
        for _ in range(3):
            comp.step()
        comp.set_values(cv0, fv0)
        torch.cuda.synchronize()
        s, e = _events()
        s.record()
        for _ in range(steps):
            comp.step()
        e.record()
        torch.cuda.synchronize()
        ms = s.elapsed_time(e)
        st = comp.stats()
        log(f"large_stencil: {name} {steps} steps in {ms:.1f} ms")
        runs[name] = {"comp": comp, "ms": ms, "mass0": m0, "mass1": comp.mass(), "stats": st}
    st = runs["f32"]["stats"]
    c32, c64 = runs["f32"]["comp"], runs["f64"]["comp"]
    dc = (c32.coarse.values.double() - c64.coarse.values)
    df = (c32.fine_field.values.double() - c64.fine_field.values)
    res.update({
        "steps": steps,
        "cells": {"coarse": st.coarse_cells, "fine": st.fine_cells, "ghost": st.ghost_cells},
        "cell_updates_per_step": st.cell_updates_per_step,
        "f32": {"ms_per_step": runs["f32"]["ms"] / steps,
                "cell_updates_per_s": st.cell_updates_per_step * steps / (runs["f32"]["ms"] * 1e-3),
                "mass_drift_rel": (runs["f32"]["mass1"] - runs["f32"]["mass0"]) / runs["f32"]["mass0"]},
        "f64": {"ms_per_step": runs["f64"]["ms"] / steps,
                "cell_updates_per_s": st.cell_updates_per_step * steps / (runs["f64"]["ms"] * 1e-3),
                "mass_drift_rel": (runs["f64"]["mass1"] - runs["f64"]["mass0"]) / runs["f64"]["mass0"]},
        "f32_vs_f64": {"coarse_max_abs": float(dc.abs().max()), "coarse_l2_rms": float(dc.pow(2).mean().sqrt()),
                       "fine_max_abs": float(df.abs().max()), "fine_l2_rms": float(df.pow(2).mean().sqrt()),
                       "mass_rel_err": abs(runs["f32"]["mass1"] - runs["f64"]["mass1"]) / abs(runs["f64"]["mass1"])},
    })
    # upwind kernel roofline on the coarse level (dense 4096^2, FP32)
    from .interval_stencil import upwind_values

    c = c32.coarse
    out = torch.empty_like(c.values)
    for _ in range(3):
        upwind_values(c, c.values, out, width=W, height=H, a=1.0, b=0.5, dt=c32.dt, dx=c32.dx_c, dy=c32.dy_c,
                      check_rows=False)
    torch.cuda.synchronize()
    s, e = _events()
    s.record()
    reps = 20
    for _ in range(reps):
        upwind_values(c, c.values, out, width=W, height=H, a=1.0, b=0.5, dt=c32.dt, dx=c32.dx_c, dy=c32.dy_c,
                      check_rows=False)
    e.record()
    torch.cuda.synchronize()
    kms = s.elapsed_time(e) / reps
    byts = 8 * c.values.numel() + 16 * c.interval_set.interval_count + 4 * c.interval_set.row_count
    res["upwind_coarse_kernel"] = {"ms": kms, "algorithmic_bytes": byts, "gbs": byts / (kms * 1e-3) / 1e9}
    f = c32.fine_field
    out = torch.empty_like(f.values)
    s, e = _events()
    s.record()
    for _ in range(reps):
        upwind_values(f, f.values, out, width=n, height=n, a=1.0, b=0.5, dt=c32.dt, dx=c32.dx_f, dy=c32.dy_f,
                      check_rows=False)
    e.record()
    torch.cuda.synchronize()
    kms = s.elapsed_time(e) / reps
    byts = 8 * f.values.numel() + 16 * f.interval_set.interval_count + 4 * f.interval_set.row_count
    res["upwind_fine_kernel"] = {"ms": kms, "algorithmic_bytes": byts, "gbs": byts / (kms * 1e-3) / 1e9}
    res["autotune"] = autotune_report(("upwind", "transfer"))
    del runs
    return res


def batched_amr(count: int, steps: int, dev, *, plain_sample: int = 16, plain_steps: int = 16,
                seed: int = 4) -> Dict:
    """Many 3-level instances through the device-regrid engine
    (subsetix_amd.amr3: each regrid + 8 steps is one hipGraph replay, no
    host synchronisation) vs the plain form (one instance at a time through
    the interval-set engine, eager launches, a host synchronisation per
    step), both counting the level-set cells advanced per step (ghost cells
    excluded).  Also: the regrid's exclusive GPU time and a step's, timed
    with events around eager launches; the timed run under torch's sync
    debug mode set to 'error' (any host read raises); FP32 masses of a sample
    against an FP64 run of the same instances (instances are independent,
    so a sample stands for itself inside the full batch)."""
    from .amr3 import DeviceBatchedAMR
    from .batch import BatchConfig, BatchedAMR

    insts = workloads.batched_instances(count=count, seed=seed)
    cfg = BatchConfig(levels=3, regrid_every=8, threshold=(0.02, 0.02), grading=1, autotune=True)
    k = cfg.regrid_every
    t0 = time.perf_counter()
    eng = DeviceBatchedAMR(insts, config=cfg, device=dev)
    torch.cuda.synchronize()
    setup_s = time.perf_counter() - t0
    log(f"batched_amr: {count} instances set up in {setup_s:.1f}s, cells {eng.cells()}")
    m0 = eng.masses().cpu().numpy()
    eng.run(2 * k)  # warm-up: both graphs captured (one per refine-mask parity) and replayed
    torch.cuda.synchronize()
    c0 = eng.cell_updates()
    with warnings.catch_warnings():
        warnings.simplefilter("ignore")  # the mode is labelled a prototype feature
        torch.cuda.set_sync_debug_mode("error")
    try:
        t0 = time.perf_counter()
        eng.run(steps)
        torch.cuda.synchronize()
        wall = time.perf_counter() - t0
    finally:
        torch.cuda.set_sync_debug_mode("default")
    cells_total = eng.cell_updates() - c0
    log(f"batched_amr: {steps} steps in {wall:.3f}s")
    m1 = eng.masses().cpu().numpy()
    total_steps = 2 * k + steps
    cells_last = eng.cells()
    # exclusive GPU times: a regrid (eager, on the settled state: it rebuilds
    # the same sets) and a step (eager launches; the state moves one step)
    s_ev, e_ev = _events()
    for _ in range(2):
        eng.regrid()
    s_ev.record()
    for _ in range(10):
        eng.regrid()
    e_ev.record()
    torch.cuda.synchronize()
    regrid_ms = s_ev.elapsed_time(e_ev) / 10
    s_ev.record()
    eng._steps(k)
    e_ev.record()
    torch.cuda.synchronize()
    step_ms = s_ev.elapsed_time(e_ev) / k
    idx = list(range(0, count, max(1, count // plain_sample)))[:plain_sample]
    m1_sample = m1[idx]
    res = {
        "engine": "device regrid: dense levels, bit-mask refine sets, per-chunk work records, one graph per "
                  "regrid + 8 steps",
        "instances": count, "steps": steps, "setup_s": setup_s, "wall_s": wall,
        "ms_per_step": wall * 1e3 / steps,
        "cells_last": cells_last, "cell_updates": cells_total,
        "cell_updates_per_s": cells_total / wall,
        "coarse_cells": int((eng.sizes.astype(np.int64) ** 2).sum()),
        "regrid_gpu_ms": regrid_ms, "step_gpu_ms": step_ms,
        "regrid_share": regrid_ms / (regrid_ms + k * step_ms),
        "host_reads_per_regrid": 0,
        "host_reads_check": "timed run under torch.cuda.set_sync_debug_mode('error'); regrid and steps issue "
                            "only asynchronous launches (no device-to-host copy in either)",
    }
    drift = np.abs((m1 - m0) / np.where(m0 != 0, m0, 1.0))
    res["mass_change_rel_max"] = float(drift.max())
    del eng
    torch.cuda.empty_cache()
    # FP64 reference of the sampled instances (same steps, same regrid cadence)
    sub = [insts[i] for i in idx]
    e64 = DeviceBatchedAMR(sub, dtype=torch.float64, config=BatchConfig(levels=3, regrid_every=8,
                                                                         threshold=(0.02, 0.02), grading=1),
                           device=dev)
    e64.run(total_steps)
    m64 = e64.masses().cpu().numpy()
    del e64
    log("batched_amr: FP64 sample done")
    rel = np.abs(m1_sample - m64) / np.abs(m64)
    res["mass_rel_err_f32_vs_f64"] = {"max": float(rel.max()), "median": float(np.median(rel)),
                                      "instances": len(idx)}
    # plain form on the same sample
    pcfg = BatchConfig(levels=3, regrid_every=8, threshold=(0.02, 0.02), grading=1, use_graph=False)
    pcells = 0
    ptime = 0.0
    for i in idx:
        e1 = BatchedAMR([insts[i]], config=pcfg, device=dev)
        e1.run(1)
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(plain_steps):
            e1.run(1)
            torch.cuda.synchronize()
            c = e1.cells
            pcells += c["level0"] + c["level1"] + c.get("level2", 0)
        ptime += time.perf_counter() - t0
        del e1
    res["plain"] = {"instances_sampled": len(idx), "steps_each": plain_steps, "cell_updates": pcells,
                    "cell_updates_per_s": pcells / ptime,
                    "desc": "one instance at a time through the interval-set engine, eager launches, host "
                            "synchronisation every step (regrids included, every 8 steps)"}
    res["batched_vs_plain"] = res["cell_updates_per_s"] / res["plain"]["cell_updates_per_s"]
    res["autotune"] = {"step": autotune_report(("upwind",)), "choices": _choice_counts(("upwind",))}
    return res


def _choice_counts(families) -> Dict:
    out = {}
    for key, e in autotune.default().table.items():
        if key[0] in families:
            d = out.setdefault(key[0], {})
            d[str(e.choice)] = d.get(str(e.choice), 0) + 1
    return out


def plain_two_level(dev, n: int = 128, steps: int = 100, regrid_every: int = 10) -> Dict:
    """Reference-shaped two-level run: 128^2 Gaussian bump (amp 1, sigma
    0.1, centre (0.3, 0.3)), refine = {u > 0.1} graded by 2 (moore), ratio 2,
    velocity (1, 0.5), CFL 0.5, FP64 field."""
    import subsetix_amd as sx
    from .amr2 import AMR2Simulation, SimulationConfig

    u0 = workloads.gaussian_bump(n)
    cfg = SimulationConfig(coarse_resolution=n, velocity=(1.0, 0.5), cfl=0.5, refine_threshold=0.1, grading=2,
                           grading_mode="moore", tag="value", accumulate="f64")
    sim = AMR2Simulation(cfg)
    f = sx.create_interval_field(sx.full_interval_set(n, n, device=dev), 0.0, dtype=torch.float64)
    f.values.copy_(torch.from_numpy(u0.reshape(-1)).to(dev))
    sim.set_initial_field(f)
    m0 = sim.total_mass()
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    sim.run(steps=steps, regrid_every=regrid_every)
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    m1 = sim.total_mass()
    st = sim.state
    refine = st.actions.refine_set()
    fine = st.geometry.fine
    cells = n * n + 4 * n * n
    return {"steps": steps, "wall_s": wall, "ms_per_step": wall * 1e3 / steps,
            "cell_updates_per_s": cells * steps / wall, "mass0": m0, "mass1": m1,
            "mass_drift_rel": (m1 - m0) / m0, "refine_intervals": refine.interval_count,
            "refine_cells": int((refine.end - refine.begin).sum().item()) if refine.interval_count else 0,
            "fine_intervals": fine.interval_count, "fine_rows": fine.row_count,
            "desc": "AMR2Simulation (full-domain levels as the reference), FP64, value tag u>=0.1, moore grading 2"}


def serve(rates, count: int, dev, seed: int = 5, budget_count: int = 400) -> Dict:
    """Served stream at each Poisson rate: continuous batching server vs the
    unbatched plain loop on the same request trace (fine domains log-uniform
    256..8192, the config-3 scale at the top).  Then a trace of config-3
    sized jobs (fine 8192^2) arriving together, large enough that admission
    by the 200 GiB budget binds: requests wait for memory to free up, and
    the queueing delay (arrival to admission) is reported apart from the
    service time (admission to result)."""
    from .serve import HBM_BUDGET_BYTES, Server, make_trace, plain_loop

    out = {"requests_per_rate": count,
           "mix": "70% advection (fine 256..8192 log-uniform on 4 sizes per octave, 4 disc layouts per size, "
                  "10..200 steps), 30% set queries (rows 256..16384), HBM budget 200 GiB", "rates": {}}
    # warm-up: kernels, graph capture path, allocator
    Server(streams=4).run(make_trace(100.0, 4, seed=seed + 100, n_fine=(128, 256), steps=(4, 8), query_rows=(64, 128)))

    def compare(trace):
        t0 = time.perf_counter()
        srv = Server(streams=4).run(trace).summary()
        t1 = time.perf_counter()
        plain = plain_loop(trace).summary()
        log(f"serve: server {t1 - t0:.1f}s, plain loop {time.perf_counter() - t1:.1f}s")
        span = max(r.arrival for r in trace)
        return {
            "server": srv, "plain": plain, "arrival_span_s": span,
            "plain_falls_behind": plain["wall_s"] > 1.05 * span,
            "cell_updates_ratio": srv["cell_updates_per_s"] / max(plain["cell_updates_per_s"], 1e-9),
            "throughput_ratio": srv["throughput_req_per_s"] / max(plain["throughput_req_per_s"], 1e-9),
        }

    for rate in rates:
        out["rates"][str(rate)] = compare(make_trace(float(rate), count, seed=seed))
    if budget_count > 0:
        trace = make_trace(1e5, budget_count, seed=seed + 7, n_fine=(8192, 8192))
        res = compare(trace)
        res["requests"] = budget_count
        res["offered_est_bytes"] = int(sum(r.est_bytes for r in trace))
        res["budget_bytes"] = HBM_BUDGET_BYTES
        out["budget_binding"] = res
    return out


def copy_bandwidth(dev, nbytes: int = 1 << 30, reps: int = 10) -> Dict:
    """Device-to-device copy of ``nbytes`` (the calibration kernel for HBM
    byte counters and the achievable-bandwidth reference)."""
    src = torch.empty(nbytes // 4, dtype=torch.float32, device=dev).fill_(1.0)
    dst = torch.empty_like(src)
    for _ in range(2):
        dst.copy_(src)
    torch.cuda.synchronize()
    s, e = _events()
    s.record()
    for _ in range(reps):
        dst.copy_(src)
    e.record()
    torch.cuda.synchronize()
    ms = s.elapsed_time(e) / reps
    return {"bytes_copied": nbytes, "ms": ms, "gbs_read_plus_write": 2 * nbytes / (ms * 1e-3) / 1e9}
