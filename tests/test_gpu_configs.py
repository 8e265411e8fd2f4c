"""GPU parity at the benchmark configurations' own workloads.

* plain two-level run: 128^2 Gaussian bump, FP64 (and FP32), 100 steps,
  regrid every 10, value tag u >= 0.1, moore grading 2, against the same
  loop on oracle primitives (simulation.py:232-354);
* large stencil + transfer: the 8192^2 seeded disc geometry (64 discs +
  annulus, 30 % coverage), FP32 and FP64 composite steps against the C
  composite step of the oracle, bit-exact; at this size every segment-table
  path runs (split-1 transfers, XCD-ordered tiles beyond the first span,
  gradient tags on 20M cells);
* batched AMR: the 512-instance draw (seed 4) run by the engine as the
  benchmark runs it (autotuned, graphs); three of its instances (the
  smallest, 64^2, the largest, 2046^2, and a median one) against the
  oracle's three-level instance engine, bit-exact, plus the nesting rules
  of test_amr_grading.py:110-248 on the whole batch;
* served stream: advection jobs drawn at the large-stencil sizes (fine
  256..8192) and set queries of 256..16384 rows, each result against the
  oracle.
"""

import numpy as np
import pytest

from oracle import oracle as ox
from tests.golden_util import assert_set_equal
from tests.gpu_util import hostset_to_oset, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


# ----------------------------------------------------------------- config 1
@pytest.mark.parametrize("dtype", ["f64", "f32"])
def test_plain_two_level_config_vs_oracle(torch_mod, dtype):
    import subsetix_amd as sx
    from subsetix_amd import workloads
    from subsetix_amd.amr2 import AMR2Simulation, SimulationConfig
    from tests.test_gpu_amr2 import _oracle_amr

    torch = torch_mod
    n = 128
    npdt = np.float64 if dtype == "f64" else np.float32
    u0 = workloads.gaussian_bump(n).reshape(-1)
    cfg = SimulationConfig(coarse_resolution=n, velocity=(1.0, 0.5), cfl=0.5, refine_threshold=0.1, grading=2,
                           grading_mode="moore", tag="value", accumulate="f64")
    sim = AMR2Simulation(cfg)
    f = sx.create_interval_field(sx.full_interval_set(n, n), 0.0,
                                 dtype=torch.float64 if dtype == "f64" else torch.float32)
    f.values.copy_(torch.from_numpy(u0.astype(npdt)))
    sim.set_initial_field(f)
    print("config 1", dtype, flush=True)
    sim.run(steps=100, regrid_every=10)
    cv, fv, refine = _oracle_amr(u0, n, 100, 10, (1.0, 0.5), 0.5, 0.1, 2, "moore", "value", "f64", npdt,
                                 return_refine=True)
    np.testing.assert_array_equal(sim.state.coarse_field.values.cpu().numpy(), cv)
    np.testing.assert_array_equal(sim.state.fine_field.values.cpu().numpy(), fv)
    got = to_host(sim.state.actions.refine_set())
    assert ox.cells(got) == ox.cells(refine)
    assert abs(sim.total_mass() - ox.mass(cv, 1.0 / n / n)) <= 1e-10 * abs(ox.mass(cv, 1.0 / n / n))


# ----------------------------------------------------------------- config 3
@pytest.fixture(scope="module")
def discs8192():
    from subsetix_amd import workloads

    discs, scale, cov = workloads.stencil_discs_config(n=8192, seed=3)
    assert abs(cov - 0.30) <= 0.01
    return hostset_to_oset(discs)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_large_stencil_config_vs_oracle(torch_mod, discs8192, dtype):
    from subsetix_amd.amr2.composite import CompositeTwoLevel
    from subsetix_amd.multilevel import restrict_set
    from tests.gpu_util import to_device

    torch = torch_mod
    tdt = torch.float32 if dtype == "f32" else torch.float64
    npdt = np.float32 if dtype == "f32" else np.float64
    W = H = 4096
    refine_d = restrict_set(to_device(discs8192), 2)
    refine = ox.restrict_set(discs8192, 2)

    assert_set_equal(to_host(refine_d), refine)
    comp = CompositeTwoLevel(W, H, refine_d, dtype=tdt, velocity=(1.0, 0.5), cfl=0.5)
    oc = ox.Composite(W, H, refine, velocity=(1.0, 0.5), cfl=0.5, dtype=npdt)
    assert_set_equal(to_host(comp.ext), oc.ext)
    assert_set_equal(to_host(comp.ghost), oc.ghost)
    rng = np.random.default_rng(3)
    cv0 = rng.random(W * H).astype(np.float32)
    fv0 = rng.random(oc.fv.size).astype(np.float32)
    comp.set_values(torch.from_numpy(cv0).cuda(), torch.from_numpy(fv0).cuda())
    oc.set_values(cv0, fv0)
    np.testing.assert_array_equal(comp.coarse.values.cpu().numpy(), oc.cv)
    np.testing.assert_array_equal(comp.fine_field.values.cpu().numpy(), oc.fv)
    comp.prepare(tune=False)
    steps = 12
    print("config 3", dtype, "steps", steps, "fine cells", oc.fv.size, flush=True)
    for _ in range(steps):
        comp.step()
        oc.step()
    np.testing.assert_array_equal(comp.coarse.values.cpu().numpy(), oc.cv)
    np.testing.assert_array_equal(comp.fine_field.values.cpu().numpy(), oc.fv)
    m, mo = comp.mass(), oc.mass()
    assert abs(m - mo) <= 1e-10 * abs(mo)  # summation orders differ


@pytest.mark.parametrize("dtype", ["f32", "f64"])
def test_large_stencil_graph_and_schedules_vs_oracle(torch_mod, discs8192, dtype):
    """Captured two-step graphs and the row-group schedule, and the gradient
    tag of the 20M-cell fine level (segment table and row groups), all
    against the oracle at the config geometry."""
    from subsetix_amd import levels
    from subsetix_amd.amr2.composite import CompositeTwoLevel
    from subsetix_amd.interval_field import IntervalField
    from subsetix_amd.interval_stencil import build_stencil_plan
    from subsetix_amd.levels import gradient_tag
    from subsetix_amd.multilevel import restrict_set
    from tests.gpu_util import to_device

    torch = torch_mod
    tdt = torch.float32 if dtype == "f32" else torch.float64
    npdt = np.float32 if dtype == "f32" else np.float64
    W = H = 4096
    refine = ox.restrict_set(discs8192, 2)
    oc = ox.Composite(W, H, refine, velocity=(-0.7, 0.4), cfl=0.5, dtype=npdt)
    rng = np.random.default_rng(8)
    cv0 = rng.random(W * H).astype(npdt)
    oc.set_values(cv0)
    for _ in range(6):
        oc.step()
    for schedule in ("graph", "rows"):
        print("config 3", dtype, schedule, flush=True)
        comp = CompositeTwoLevel(W, H, restrict_set(to_device(discs8192), 2), dtype=tdt, velocity=(-0.7, 0.4),
                                 cfl=0.5)
        comp.set_values(torch.from_numpy(cv0).cuda())
        if schedule == "graph":
            s = torch.cuda.Stream()
            g = comp.capture(s)
            for _ in range(3):
                g.replay()
            torch.cuda.synchronize()
        else:
            from subsetix_amd import interval_stencil

            levels.set_transfer_schedule("rows")
            interval_stencil.set_stencil_schedule("rows")
            try:
                for _ in range(6):
                    comp.step()
            finally:
                levels.set_transfer_schedule("auto")
                interval_stencil.set_stencil_schedule("auto")
        np.testing.assert_array_equal(comp.coarse.values.cpu().numpy(), oc.cv)
        np.testing.assert_array_equal(comp.fine_field.values.cpu().numpy(), oc.fv)
    f = comp.fine_field
    fv = IntervalField(f.interval_set, f.values.clone(), f.interval_cell_offsets)
    plan = build_stencil_plan(fv.interval_set, fv.interval_cell_offsets, 2 * W)
    for t in (0.05, 0.2):
        exp = ox.grad_tag(oc.ext, oc.fv, t)
        for p in (plan, None):
            print("config 3 gradient tag", dtype, t, "seg" if p is not None else "rows", flush=True)
            assert_set_equal(to_host(gradient_tag(fv, t, plan=p)), exp)


# ----------------------------------------------------------------- config 4
def _slice(s: ox.OSet, vals: np.ndarray, base: int, nrows: int):
    """The rows [base, base + nrows) of a stacked set (row ids shifted to 0)
    and their values."""
    r0 = int(np.searchsorted(s.rows, base))
    r1 = int(np.searchsorted(s.rows, base + nrows))
    i0, i1 = int(s.off[r0]), int(s.off[r1])
    co = ox.cell_offsets(s)
    sub = ox.OSet((s.rows[r0:r1] - base).astype(np.int32), (s.off[r0:r1 + 1] - i0).astype(np.int32),
                  s.beg[i0:i1].copy(), s.end[i0:i1].copy())
    v = None if vals is None else vals[co[i0]:co[i1]]
    return sub, v


def _compare_instance(eng, i, inst, what):
    n = int(eng.sizes[i])
    for lv in range(eng.cfg.levels):
        base = int(eng.bases[lv][i].item())
        nr = n * 2 ** lv
        ext, vals = _slice(to_host(eng.ext[lv]), eng.buf[lv][eng.cur].cpu().numpy(), base, nr)
        exp_ext = ox.compact_rows(inst.ext[lv])
        assert_set_equal(ox.compact_rows(ext), exp_ext)
        np.testing.assert_array_equal(vals, inst.vals[lv], err_msg=f"{what} level {lv}")
        lvset, _ = _slice(to_host(eng.level_sets[lv]), None, base, nr)
        assert_set_equal(ox.compact_rows(lvset), ox.compact_rows(inst.sets[lv]))


def _picks(insts):
    ns = np.array([i["n"] for i in insts])
    med = int(np.argsort(ns)[ns.size // 2])
    return [int(np.argmin(ns)), med, int(np.argmax(ns))]


def _oracle_instance(eng, i, insts, dtype, cfg):
    n = int(eng.sizes[i])
    cs = eng.cstart0.cpu().numpy()
    return ox.Instance3(n, eng.level0_values()[int(cs[i]):int(cs[i + 1])].cpu().numpy(), insts[i]["velocity"],
                        dtype=dtype, levels=cfg.levels, cfl=cfg.cfl, threshold=cfg.threshold, grading=cfg.grading,
                        regrid_every=cfg.regrid_every, accumulate=cfg.accumulate)


def test_batched_amr_config_vs_oracle(torch_mod):
    """The 512-instance draw, autotuned and graph-captured as benchmarked,
    10 steps (regrids before steps 0 and 8); three instances against the
    oracle; nesting rules on every instance."""
    from subsetix_amd import workloads
    from subsetix_amd.batch import BatchConfig, BatchedAMR

    insts = workloads.batched_instances(count=512, seed=4)
    cfg = BatchConfig(levels=3, regrid_every=8, threshold=(0.02, 0.02), grading=1, autotune=True)
    eng = BatchedAMR(insts, config=cfg)
    picks = _picks(insts)
    assert [insts[i]["n"] for i in (picks[0], picks[2])] == [64, 2046]
    oracles = [_oracle_instance(eng, i, insts, np.float32, cfg) for i in picks]
    print("config 4: engine set up", eng.cells, flush=True)
    eng.run(10)
    for o in oracles:
        o.run(10)
    for i, o in zip(picks, oracles):
        print("config 4 instance", i, insts[i]["n"], flush=True)
        _compare_instance(eng, i, o, f"instance {i}")
    # nesting on the whole batch (oracle set algebra on the device sets)
    l0, l1, l2 = (to_host(s) for s in eng.level_sets)
    D1 = to_host(eng.domains[1])
    r0, r1 = (to_host(s) for s in eng.refines)
    parents = ox.restrict_set(l2, 2)
    assert ox.difference(parents, l1).n == 0  # level 2 inside level 1
    contact = ox.intersection(ox.dilate(parents, 1, 1), D1)
    assert ox.difference(contact, l1).n == 0  # no 8-neighbour contact with cells outside level 1
    assert ox.difference(ox.restrict_set(l1, 2), r0).n == 0 and ox.difference(r0, ox.restrict_set(l1, 2)).n == 0
    assert ox.difference(ox.prolong_set(r1, 2), l2).n == 0 and ox.difference(l2, ox.prolong_set(r1, 2)).n == 0
    m = eng.masses().cpu().numpy()
    for i, o in zip(picks, oracles):
        assert abs(m[i] - o.mass()) <= 1e-10 * abs(o.mass())


def test_batched_amr_fp64_instances_vs_oracle(torch_mod):
    from subsetix_amd import workloads
    from subsetix_amd.batch import BatchConfig, BatchedAMR

    torch = torch_mod
    insts = workloads.batched_instances(count=512, seed=4)
    picks = _picks(insts)[:2] + [int(np.argsort([i["n"] for i in insts])[400])]
    sub = [insts[i] for i in picks]
    cfg = BatchConfig(levels=3, regrid_every=8, threshold=(0.02, 0.02), grading=1)
    eng = BatchedAMR(sub, dtype=torch.float64, config=cfg)
    oracles = [_oracle_instance(eng, j, sub, np.float64, cfg) for j in range(len(sub))]
    eng.run(10)
    for j, o in enumerate(oracles):
        print("config 4 fp64 instance", picks[j], sub[j]["n"], flush=True)
        o.run(10)
        _compare_instance(eng, j, o, f"fp64 instance {picks[j]}")


# ----------------------------------------------------------------- config 5
def test_served_stream_config_vs_oracle(torch_mod):
    """Advection jobs at the large-stencil sizes and queries of 256..16384
    rows through the continuous-batching server; every result against the
    oracle (composite steps from the same seeded initial values)."""
    from subsetix_amd.serve import Server, make_trace

    torch = torch_mod
    trace = make_trace(2000.0, 10, seed=21, n_fine=(256, 8192), steps=(10, 16), query_rows=(256, 16384))
    assert max(r.payload["n_fine"] for r in trace if r.kind == "advect") > 4096
    res = Server(streams=3, chunk=4, keep_outputs=True).run(trace).results
    dev = torch.device("cuda", torch.cuda.current_device())
    for r in trace:
        out = res[r.rid].output
        print("config 5 request", r.rid, r.kind, r.payload.get("n_fine", r.payload.get("rows")), flush=True)
        if r.kind == "query":
            A = ox.make(r.payload["a"]["off"], r.payload["a"]["begin"], r.payload["a"]["end"])
            B = ox.make(r.payload["b"]["off"], r.payload["b"]["begin"], r.payload["b"]["end"])
            for op in ("union", "intersection", "difference"):
                exp = ox.setop(op, A, B)
                off, b, e = out[op]
                np.testing.assert_array_equal(off, exp.off)
                np.testing.assert_array_equal(b, exp.beg)
                np.testing.assert_array_equal(e, exp.end)
            continue
        p = r.payload
        nc = p["n_fine"] // 2
        d = p["discs"]
        oc = ox.Composite(nc, nc, ox.restrict_set(ox.make(d["off"], d["begin"], d["end"], d["rows"]), 2),
                          velocity=p["velocity"], cfl=0.5, dtype=np.float32)
        gen = torch.Generator(device=dev)
        gen.manual_seed(int(p["seed"]))
        cv = torch.rand(nc * nc, generator=gen, device=dev, dtype=torch.float32).cpu().numpy()
        fv = torch.rand(oc.fv.size, generator=gen, device=dev, dtype=torch.float32).cpu().numpy()
        oc.set_values(cv, fv)
        for _ in range(int(p["steps"])):
            oc.step()
        np.testing.assert_array_equal(out["coarse"], oc.cv)
        np.testing.assert_array_equal(out["fine"], oc.fv)
        assert abs(out["mass"] - oc.mass()) <= 1e-10 * abs(oc.mass())
