"""GPU: the batched multi-instance AMR engine.

* batch of many == each instance alone (bit-exact, every level);
* hipGraph replay == eager launches (bit-exact);
* the batched upwind equals the single-instance kernel on each instance;
* 3-level nesting invariants after regrid (test_amr_grading.py:110-248):
  L2 inside L1, no 8-neighbour contact of L2 with cells outside L1, coarse
  parents of L2 refined at L1, block-uniform levels.
"""

import numpy as np
import pytest

from oracle import oracle as ox
from tests.gpu_util import to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


def _instances():
    from subsetix_amd import workloads

    inst = workloads.batched_instances(count=5, seed=11, nmin=16, nmax=72)
    return inst


def _level_cells(eng, lv, idx):
    """{instance: {(y, x): value}} of level lv (level sets only, no ghosts)."""
    ext = to_host(eng.ext[lv])
    vals = eng.buf[lv][idx].cpu().numpy()
    co = ox.cell_offsets(ext)
    ir = ox.interval_rows(ext)
    lvset = ox.cells(to_host(eng.level_sets[lv]))
    bases = eng.bases[lv].cpu().numpy()
    out = {}
    for i in range(ext.n):
        y = int(ir[i])
        inst = int(np.searchsorted(bases, y, side="right") - 1)
        for k in range(int(ext.end[i] - ext.beg[i])):
            x = int(ext.beg[i]) + k
            if (y, x) in lvset:
                out.setdefault(inst, {})[(y - int(bases[inst]), x)] = vals[co[i] + k]
    return out


def test_batch_equals_alone_and_graph_equals_eager(torch_mod):
    from subsetix_amd.batch import BatchConfig, BatchedAMR

    insts = _instances()
    cfg = BatchConfig(levels=3, regrid_every=4, threshold=(0.02, 0.02), grading=1)
    many = BatchedAMR(insts, config=cfg)
    many.run(10)
    eager_cfg = BatchConfig(levels=3, regrid_every=4, threshold=(0.02, 0.02), grading=1, use_graph=False)
    eager = BatchedAMR(insts, config=eager_cfg)
    eager.run(10)
    for lv in range(3):
        a = _level_cells(many, lv, many.cur)
        b = _level_cells(eager, lv, eager.cur)
        assert a.keys() == b.keys()
        for i in a:
            assert a[i] == b[i], (lv, i)
    np.testing.assert_array_equal(many.masses().cpu().numpy(), eager.masses().cpu().numpy())
    for k, inst in enumerate(insts):
        alone = BatchedAMR([inst], config=cfg)
        alone.run(10)
        for lv in range(3):
            a = _level_cells(many, lv, many.cur).get(k, {})
            b = _level_cells(alone, lv, alone.cur).get(0, {})
            assert a == b, (k, lv)
        assert many.masses().cpu().numpy()[k] == alone.masses().cpu().numpy()[0]


@pytest.mark.parametrize("dt_name", ["f32", "f64"])
def test_batched_upwind_matches_single_instance(torch_mod, dt_name):
    """Stacked instances of sizes 16..72 (instance boundaries fall inside
    waves, so lanes of one wave read two instances' coefficients): the
    segment and row-group kernels agree, and each instance equals its
    single-instance step."""
    torch = torch_mod
    import subsetix_amd as sx
    from subsetix_amd.batch import BatchConfig, BatchedAMR
    from subsetix_amd.interval_stencil import upwind_values

    dtype = torch.float32 if dt_name == "f32" else torch.float64
    insts = _instances()
    eng = BatchedAMR(insts, dtype=dtype, config=BatchConfig(levels=1, use_graph=False))
    f = eng._fields(eng.cur)[0]
    out = torch.empty_like(f.values)
    from subsetix_amd.levels import upwind_batched

    upwind_batched(f, f.values, out, eng.row_inst[0], eng.inst_w[0], eng.inst_coef[0])
    out_t = torch.empty_like(f.values)
    upwind_batched(f, f.values, out_t, eng.row_inst[0], eng.inst_w[0], eng.inst_coef[0], plan=eng.plans[0])
    np.testing.assert_array_equal(out_t.cpu().numpy(), out.cpu().numpy())
    off = 0
    for k, (inst, n) in enumerate(zip(insts, eng.sizes)):
        n = int(n)
        s = sx.full_interval_set(n, n)
        fi = sx.create_interval_field(s, 0.0, dtype=dtype)
        fi.values.copy_(f.values[off:off + n * n])
        o = torch.empty_like(fi.values)
        a, b = inst["velocity"]
        dt = float(eng.inst_coef[0][5 * k + 2])  # the coefficient (of the field's type) the batched kernel uses
        upwind_values(fi, fi.values, o, width=n, height=n, a=a, b=b, dt=dt, dx=1.0 / n, dy=1.0 / n)
        np.testing.assert_array_equal(o.cpu().numpy(), out[off:off + n * n].cpu().numpy())
        off += n * n


def test_stacked_level0_gradient_tag_segments_equal_rows(torch_mod):
    """The regrid tag of the stacked level 0 through the level-0 segment
    table (per-instance widths) equals the row-group kernel."""
    from subsetix_amd.batch import BatchConfig, BatchedAMR
    from subsetix_amd.levels import gradient_tag

    eng = BatchedAMR(_instances(), config=BatchConfig(levels=1, use_graph=False))
    eng.run(3)
    f = eng._fields(eng.cur)[0]
    for t in (0.005, 0.02):
        a = to_host(gradient_tag(f, t, plan=eng._plan0))
        b = to_host(gradient_tag(f, t))
        for name in ("rows", "off", "beg", "end"):
            np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=name)
        assert a.n > 0


def _dilate8(cells):
    out = set(cells)
    for (y, x) in cells:
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                out.add((y + dy, x + dx))
    return out


def test_three_level_nesting_invariants(torch_mod):
    from subsetix_amd.batch import BatchConfig, BatchedAMR

    insts = _instances()
    eng = BatchedAMR(insts, config=BatchConfig(levels=3, regrid_every=3, threshold=(0.015, 0.015), grading=1))
    for _ in range(3):
        eng.run(3)
        l1 = ox.cells(to_host(eng.level_sets[1]))
        l2 = ox.cells(to_host(eng.level_sets[2]))
        r0 = ox.cells(to_host(eng.refines[0]))
        r1 = ox.cells(to_host(eng.refines[1]))
        assert l1 and l2, "expected refinement on both levels"
        l1_fine = {(2 * y + dy, 2 * x + dx) for (y, x) in l1 for dy in (0, 1) for dx in (0, 1)}
        assert l2 <= l1_fine                                   # L2 inside L1
        dom2 = ox.cells(to_host(eng.domains[2]))
        assert (_dilate8(l2) & dom2) <= l1_fine                # no 8-contact with outside L1
        assert {(y // 2, x // 2) for (y, x) in l2} <= l1       # parents of L2 refined at L1
        # block uniformity: levels are prolongations of their refine sets
        assert l1 == {(2 * y + dy, 2 * x + dx) for (y, x) in r0 for dy in (0, 1) for dx in (0, 1)}
        assert l2 == {(2 * y + dy, 2 * x + dx) for (y, x) in r1 for dy in (0, 1) for dx in (0, 1)}
