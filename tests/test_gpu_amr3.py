"""GPU: the batched three-level engine with the device-only regrid
(subsetix_amd.amr3.DeviceBatchedAMR) against the oracle's three-level
instance engine, bit-exact: level sets, level values (on the level sets),
level-0 values and masses; captured regrid+steps graphs == eager launches;
the 512-instance draw of the batched benchmark."""

import numpy as np
import pytest

from oracle import oracle as ox

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _dense(s: ox.OSet, vals, side, dtype=np.float64):
    """Dense (side, side) mask and values of a set / field of one instance."""
    mask = np.zeros((side, side), bool)
    out = np.zeros((side, side), dtype)
    co = ox.cell_offsets(s)
    ir = ox.interval_rows(s)
    for k in range(s.n):
        y, b, e = int(ir[k]), int(s.beg[k]), int(s.end[k])
        mask[y, b:e] = True
        if vals is not None:
            out[y, b:e] = vals[co[k]:co[k + 1]]
    return mask, out


def _compare(eng, j, o, what):
    n = int(eng.sizes[j])
    for lv in range(3):
        side = n << lv
        got_mask = eng.level_mask(lv, j)
        exp_mask, _ = _dense(o.sets[lv], None, side)
        np.testing.assert_array_equal(got_mask, exp_mask, err_msg=f"{what} level {lv} set")
        _, exp_vals = _dense(o.ext[lv], o.vals[lv], side, dtype=o.vals[lv].dtype)
        got = eng.level_values(lv, j)
        np.testing.assert_array_equal(got[exp_mask], exp_vals[exp_mask], err_msg=f"{what} level {lv} values")


def _oracles(eng, insts, dtype, cfg, idx):
    return [ox.Instance3(int(insts[i]["n"]), eng.level_values(0, i).reshape(-1), insts[i]["velocity"], dtype=dtype,
                         levels=3, cfl=cfg.cfl, threshold=cfg.threshold, grading=cfg.grading,
                         regrid_every=cfg.regrid_every, accumulate=cfg.accumulate) for i in idx]


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_device_regrid_engine_vs_oracle(torch_mod, dt):
    torch = torch_mod
    from subsetix_amd import workloads
    from subsetix_amd.amr3 import DeviceBatchedAMR
    from subsetix_amd.batch import BatchConfig

    insts = workloads.batched_instances(count=7, seed=11, nmin=16, nmax=90)
    tdt = torch.float32 if dt == "f32" else torch.float64
    npdt = np.float32 if dt == "f32" else np.float64
    cfg = BatchConfig(levels=3, regrid_every=4, threshold=(0.02, 0.02), grading=1, use_graph=False)
    eng = DeviceBatchedAMR(insts, dtype=tdt, config=cfg)
    oracles = _oracles(eng, insts, npdt, cfg, range(len(insts)))
    for j, o in enumerate(oracles):
        _compare(eng, j, o, f"{dt} instance {j} after the first regrid")
    for steps in (3, 6, 5):  # crosses regrids before steps 4, 8, 12
        print("device regrid engine", dt, "steps", steps, flush=True)
        eng.run(steps)
        for o in oracles:
            o.run(steps)
        for j, o in enumerate(oracles):
            _compare(eng, j, o, f"{dt} instance {j} step {eng.step_count}")
    assert any(o.sets[2].n for o in oracles), "no instance reached level 2"
    m = eng.masses().cpu().numpy()
    for j, o in enumerate(oracles):
        assert abs(m[j] - o.mass()) <= 1e-10 * abs(o.mass())
    cells = [sum(int(ox.cell_offsets(s)[-1]) for s in o.sets) for o in oracles]
    c = eng.cells()
    assert c["level0"] + c["level1"] + c["level2"] == sum(cells)
    eng.check_status()


def test_device_regrid_graph_equals_eager(torch_mod):
    torch = torch_mod
    from subsetix_amd import workloads
    from subsetix_amd.amr3 import DeviceBatchedAMR
    from subsetix_amd.batch import BatchConfig

    insts = workloads.batched_instances(count=9, seed=21, nmin=24, nmax=130)
    outs = []
    for use_graph in (False, True):
        cfg = BatchConfig(levels=3, regrid_every=8, threshold=(0.02, 0.02), grading=1, use_graph=use_graph)
        eng = DeviceBatchedAMR(insts, config=cfg)
        eng.run(8)   # first cycle: steps only (the construction regridded)
        eng.run(20)  # two full cycles (graph replays when captured) and four steps
        torch.cuda.synchronize()
        outs.append(eng)
    a, b = outs
    assert b._graphs, "no graph was captured"
    for lv in range(3):
        for j in range(len(insts)):
            ma, mb = a.level_mask(lv, j), b.level_mask(lv, j)
            np.testing.assert_array_equal(ma, mb)
            np.testing.assert_array_equal(a.level_values(lv, j)[ma], b.level_values(lv, j)[mb])
    assert a.cell_updates() == b.cell_updates() > 0
    np.testing.assert_array_equal(a.masses().cpu().numpy(), b.masses().cpu().numpy())


def test_device_regrid_config_batch_vs_oracle(torch_mod):
    """The 512-instance draw with graph-captured regrid cycles; the smallest,
    a median and the largest instance against the oracle after 18 steps."""
    from subsetix_amd import workloads
    from subsetix_amd.amr3 import DeviceBatchedAMR
    from subsetix_amd.batch import BatchConfig

    insts = workloads.batched_instances(count=512, seed=4)
    ns = np.array([i["n"] for i in insts])
    picks = [int(np.argmin(ns)), int(np.argsort(ns)[ns.size // 2]), int(np.argmax(ns))]
    cfg = BatchConfig(levels=3, regrid_every=8, threshold=(0.02, 0.02), grading=1)
    eng = DeviceBatchedAMR(insts, config=cfg)
    oracles = _oracles(eng, insts, np.float32, cfg, picks)
    print("config batch: engine set up", eng.cells(), flush=True)
    eng.run(18)
    for o in oracles:
        o.run(18)
    for i, o in zip(picks, oracles):
        print("config batch instance", i, insts[i]["n"], flush=True)
        # level_mask / level_values take the instance index in the batch
        for lv in range(3):
            side = int(insts[i]["n"]) << lv
            exp_mask, exp_vals = _dense(o.sets[lv], None, side)
            np.testing.assert_array_equal(eng.level_mask(lv, i), exp_mask)
            _, exp_vals = _dense(o.ext[lv], o.vals[lv], side, dtype=np.float32)
            np.testing.assert_array_equal(eng.level_values(lv, i)[exp_mask], exp_vals[exp_mask])
    m = eng.masses().cpu().numpy()
    for i, o in zip(picks, oracles):
        assert abs(m[i] - o.mass()) <= 1e-10 * abs(o.mass())
    eng.check_status()


@pytest.mark.parametrize("grading,accumulate", [(2, "f64"), (1, "f32")])
def test_device_regrid_edge_sizes_vs_oracle(torch_mod, grading, accumulate):
    """Sizes at the edges of the chunk layout (1, 2, 3, 31, 32, 33, 63, 64,
    65, 127 coarse cells: level rows of 1..508 cells, partial and exact
    64-cell chunks), a zero velocity component, grading 2 and FP32
    restriction sums, against the oracle's instance engine."""
    from subsetix_amd import workloads
    from subsetix_amd.amr3 import DeviceBatchedAMR
    from subsetix_amd.batch import BatchConfig

    base = workloads.batched_instances(count=10, seed=33, nmin=16, nmax=64)
    sizes = [1, 2, 3, 31, 32, 33, 63, 64, 65, 127]
    insts = [dict(b, n=n) for b, n in zip(base, sizes)]
    insts[3] = dict(insts[3], velocity=(0.0, insts[3]["velocity"][1]))
    insts[7] = dict(insts[7], velocity=(insts[7]["velocity"][0], 0.0))
    cfg = BatchConfig(levels=3, regrid_every=3, threshold=(0.02, 0.02), grading=grading, accumulate=accumulate,
                      use_graph=True)
    eng = DeviceBatchedAMR(insts, config=cfg)
    oracles = _oracles(eng, insts, np.float32, cfg, range(len(insts)))
    for steps in (2, 5, 4):
        print("edge sizes", grading, accumulate, "steps", steps, flush=True)
        eng.run(steps)
        for o in oracles:
            o.run(steps)
        for j, o in enumerate(oracles):
            _compare(eng, j, o, f"grading {grading} {accumulate} instance {j} (n {sizes[j]}) step {eng.step_count}")
    m = eng.masses().cpu().numpy()
    for j, o in enumerate(oracles):
        assert abs(m[j] - o.mass()) <= 1e-10 * max(abs(o.mass()), 1e-300)
    eng.check_status()
