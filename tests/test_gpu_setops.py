"""GPU parity: set algebra (fused union / intersection / difference kernel)
against the oracle and the reference's golden vectors, bit-exact."""

import numpy as np
import pytest

from oracle import oracle as ox
from tests.golden_util import assert_set_equal, cell_model, eval_spec, oset, random_canonical
from tests.gpu_util import hostset_to_oset, to_device, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import subsetix_amd

    return subsetix_amd


def _ops(sx):
    def wrap(fn):
        return lambda a, b: sx.evaluate(fn(sx.make_input(a), sx.make_input(b)))

    return {
        "union": wrap(sx.make_union),
        "intersection": wrap(sx.make_intersection),
        "difference": wrap(sx.make_difference),
        "symmetric_difference": wrap(sx.make_symmetric_difference),
        "complement": wrap(sx.make_complement),
    }


def test_golden_expressions(sx, golden):
    surf = {k: to_device(oset(v)) for k, v in golden["surfaces"].items()}
    ops = _ops(sx)
    for case in golden["expressions"]:
        print("case", case["id"], flush=True)
        got = eval_spec(case["spec"], surf, ops)
        assert_set_equal(to_host(got), oset(case["expect"]))


def test_golden_cell_model(sx, golden):
    surf_h = {k: oset(v) for k, v in golden["surfaces"].items()}
    surf = {k: to_device(v) for k, v in surf_h.items()}
    cells = {k: ox.cells(v) for k, v in surf_h.items()}
    ops = _ops(sx)
    for group in ("dense", "sparse"):
        for spec in golden["cell_model_specs"][group]:
            got = eval_spec(spec, surf, ops)
            assert ox.cells(to_host(got)) == cell_model(spec, cells), spec


@pytest.mark.parametrize("group", [None, -2, -3, -4, -5, -6, -7, -8, 4, 16, 64])  # None / negative: LDS tile shapes
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_random_vs_oracle(sx, seed, group):
    from subsetix_amd import _core

    rng = np.random.default_rng(seed)
    _core.set_group_override(group)
    try:
        for rows_a, rows_b, touching in (
            (None, None, False),
            (np.arange(0, 40, 2), np.arange(0, 40, 3), False),
            (np.arange(-5, 30), np.arange(10, 12), True),
        ):
            na = 40 if rows_a is None else rows_a.size
            nb = 40 if rows_b is None else rows_b.size
            a = random_canonical(rng, na, 400, mean_n=int(rng.integers(1, 80)), rows=rows_a, touching=touching)
            b = random_canonical(rng, nb, 400, mean_n=int(rng.integers(1, 80)), rows=rows_b, touching=touching)
            da, db = to_device(a), to_device(b)
            for op in ("union", "intersection", "difference"):
                print("seed", seed, "group", group, "op", op, "rows", na, nb, flush=True)
                got = to_host(_ops(sx)[op](da, db))
                assert_set_equal(got, ox.setop(op, a, b))
    finally:
        _core.set_group_override(None)


def test_empty_and_degenerate(sx):
    dev_empty = to_device(ox.make([0, 0, 0], [], []))
    a = to_device(ox.make([0, 2, 3], [0, 6, 1], [4, 9, 5]))
    ops = _ops(sx)
    for op in ("union", "intersection", "difference"):
        got = to_host(ops[op](dev_empty, dev_empty))
        assert got.n == 0 and list(got.off) == [0, 0, 0]
        assert_set_equal(to_host(ops[op](a, dev_empty)), ox.setop(op, to_host(a), to_host(dev_empty)))
        assert_set_equal(to_host(ops[op](dev_empty, a)), ox.setop(op, to_host(dev_empty), to_host(a)))
    zero_rows = to_device(ox.make([0], [], []))
    got = to_host(ops["union"](zero_rows, zero_rows))
    assert got.nrows == 0 and list(got.off) == [0]


def test_setalgebra_config_full_size(sx):
    """The set-algebra microbenchmark inputs (4096 rows x 65536, Poisson(32)
    intervals per row) at full size, bit-exact."""
    from subsetix_amd import workloads

    ha, hb = workloads.setalgebra_pair()
    a, b = hostset_to_oset(ha), hostset_to_oset(hb)
    da, db = to_device(a), to_device(b)
    for op in ("union", "intersection", "difference"):
        print("4096x65536 pair", op, flush=True)
        assert_set_equal(to_host(_ops(sx)[op](da, db)), ox.setop(op, a, b))


def test_workspace_reuses_output_buffers(sx, golden):
    """test_expressions.py:270-284: re-evaluating with a workspace reuses the
    begin buffer."""
    a = to_device(oset(golden["surfaces"]["A"]))
    b = to_device(oset(golden["surfaces"]["B"]))
    ws = sx.CuPyWorkspace()
    expr = sx.make_union(sx.make_input(a), sx.make_input(b))
    r1 = sx.evaluate(expr, ws)
    p1 = r1.begin.data_ptr()
    del r1
    r2 = sx.evaluate(expr, ws)
    assert r2.begin.data_ptr() == p1
    # nested expressions with a workspace stay correct (no shared offsets)
    nested = sx.make_union(sx.make_difference(sx.make_input(a), sx.make_input(b)),
                           sx.make_intersection(sx.make_input(a), sx.make_input(b)))
    got = to_host(sx.evaluate(nested, ws))
    ha, hb = to_host(a), to_host(b)
    assert_set_equal(got, ox.union(ox.difference(ha, hb), ox.intersection(ha, hb)))


def test_invalid_inputs_raise(sx):
    import torch

    with pytest.raises(TypeError):
        sx.IntervalSet(begin=torch.zeros(1, dtype=torch.int64), end=torch.zeros(1, dtype=torch.int32),
                       row_offsets=torch.tensor([0, 1], dtype=torch.int32), rows=torch.tensor([0], dtype=torch.int32))
    with pytest.raises(ValueError):
        sx.build_interval_set([0, 1, 2], [0, 1], [1, 2], rows=[3, 3])
    with pytest.raises(ValueError):
        sx.build_interval_set([0, 2], [0], [1])


def test_tile_fallback_long_rows(sx):
    """Rows whose intervals overflow the LDS tile (more than 2048 merged
    intervals) take the per-row wavefront path inside the same launch."""
    rng = np.random.default_rng(42)
    for counts in ([3000, 2, 0, 5], [1, 1500, 1500, 0, 7, 2600], [0] * 30 + [5000]):
        off, beg, end = [0], [], []
        offb, begb, endb = [0], [], []
        for c in counts:
            x = 0
            for _ in range(c):
                b = x + int(rng.integers(1, 4))
                e = b + int(rng.integers(1, 4))
                beg.append(b)
                end.append(e)
                x = e
            off.append(len(beg))
            x = 0
            for _ in range(c // 2 + 1):
                b = x + int(rng.integers(1, 6))
                e = b + int(rng.integers(1, 6))
                begb.append(b)
                endb.append(e)
                x = e
            offb.append(len(begb))
        a = ox.make(off, beg, end)
        b = ox.make(offb, begb, endb)
        da, db = to_device(a), to_device(b)
        for op in ("union", "intersection", "difference"):
            print("long rows", counts[:3], op, flush=True)
            assert_set_equal(to_host(_ops(sx)[op](da, db)), ox.setop(op, a, b))


def test_row_count_multiple_of_tile_rows(sx):
    """Row counts that the tile height divides (the launch then ends with an
    empty tile): exact-capacity merges and binary operations stay exact.
    Tile height = min(256, floor(1536 / mean intervals per row))."""
    rng = np.random.default_rng(7)
    for nrows, per_row in ((256, 1), (512, 1), (384, 4), (768, 2)):
        off = np.arange(nrows + 1, dtype=np.int64) * per_row
        beg, end = [], []
        for _ in range(nrows):
            x = int(rng.integers(0, 5))
            for _ in range(per_row):
                b = x + int(rng.integers(2, 9))
                e = b + int(rng.integers(1, 9))
                beg.append(b)
                end.append(e)
                x = e
        a = ox.make(off.tolist(), beg, end)
        b = ox.translate(a, 1, 0)
        da, db = to_device(a), to_device(b)
        print("rows", nrows, "per row", per_row, flush=True)
        assert_set_equal(to_host(sx.dilate_interval_set(da, 1, 0)), ox.dilate(a, 1, 0))
        for op in ("union", "intersection", "difference"):
            assert_set_equal(to_host(_ops(sx)[op](da, db)), ox.setop(op, a, b))


@pytest.mark.parametrize("schedule", [-1, -2, -3, -4, -5, -6, -7, -8])
def test_tile_schedules_batched_pairs(sx, schedule):
    """Tile schedules on 16 stacked full-size pairs (hundreds of tiles),
    bit-exact against the oracle: single pass 512 x 8 (-1, decoupled
    look-back), two passes (-2, count kernel, prefix kernel, write kernel),
    single pass 128 x 16 (-3) and 64 x 32 (-4), flat tiles 512 x 8 (-5) and
    1024 x 8 (-6), flat tiles in the register-output form (-7) and the same with count-balanced tiles (-8);
    repeated launches on one scratch buffer alternate schedules.  Dilation (transforms on load, one operand) runs
    through the same shape (-7 and -8 take the -6 tiles there)."""
    import torch

    from subsetix_amd import _core, workloads

    pairs = [workloads.setalgebra_pair(seed_a=11 + 2 * k, seed_b=12 + 2 * k) for k in range(16)]
    A = workloads.stack_rows([p[0] for p in pairs])
    B = workloads.stack_rows([p[1] for p in pairs])
    a, b = hostset_to_oset(A), hostset_to_oset(B)
    da, db = to_device(a), to_device(b)
    R = da.row_count
    for op, name in ((0, "union"), (1, "intersection"), (2, "difference")):
        exp = ox.setop(name, a, b)
        for g in (schedule, -2 if schedule == -1 else -1, schedule):
            print("schedule", g, name, flush=True)
            out_off, ob, oe, ctx = _core.setop_raw(op, (da.begin, da.end, da.row_offsets),
                                                   (db.begin, db.end, db.row_offsets), nrows=R, group=g)
            (total,) = _core.read_counts(ctx, ctx.words[1])
            np.testing.assert_array_equal(out_off.cpu().numpy(), exp.off)
            np.testing.assert_array_equal(ob[:total].cpu().numpy(), exp.beg)
            np.testing.assert_array_equal(oe[:total].cpu().numpy(), exp.end)
    small_a = hostset_to_oset(pairs[0][0])
    _core.set_group_override(schedule)
    try:
        for r in (1, 3):
            print("schedule", schedule, "dilate", r, flush=True)
            assert_set_equal(to_host(sx.dilate_interval_set(to_device(small_a), r, r)), ox.dilate(small_a, r, r))
    finally:
        _core.set_group_override(None)


def _row_runs(rng, base, n, touching):
    out, x = [], base + int(rng.integers(0, 4))
    for _ in range(n):
        gap = int(rng.geometric(0.3)) - (1 if touching and rng.random() < 0.3 else 0)
        b = x + max(gap, 0 if touching else 1)
        e = b + int(rng.geometric(0.25))
        out.append((b, e))
        x = e
    return out


def _rows_to_oset(rows):
    off, beg, end = [0], [], []
    for lst in rows:
        for b, e in lst:
            beg.append(b)
            end.append(e)
        off.append(len(beg))
    return ox.make(off, beg, end)


def _flat_edge_pair(seed):
    """Identity-row pair (4000 rows) whose tiles reach every branch of the
    flat tile kernel: flat tiles with negative and with large coordinates,
    tiles whose row bands overflow 31 bits (coordinates near +-1e9 and near
    the int32 limits: per-row fallback), a row longer than the LDS capacity
    of every tile shape (per-row fallback), runs of empty rows (900 of them:
    whole tiles without an interval), rows present in one operand only,
    touching runs."""
    rng = np.random.default_rng(seed)
    A, B = [], []
    for r in range(4000):
        base = 0
        if 200 <= r < 260:
            base = -5000
        elif 700 <= r < 704:
            base = 1_000_000_000 + 40_000_000 * (r - 700)
        elif 1500 <= r < 1502:
            base = -2_000_000_000
        elif 1800 <= r < 1803:
            base = 3_000_000
        la = _row_runs(rng, base, int(rng.poisson(12)), r % 7 == 0)
        lb = _row_runs(rng, base + int(rng.integers(-20, 20)), int(rng.poisson(12)), r % 5 == 0)
        if 300 <= r < 360 or 2500 <= r < 3400:
            la, lb = [], []
        elif 400 <= r < 420:
            lb = []
        elif 420 <= r < 440:
            la = []
        elif r == 1000:
            la = _row_runs(rng, 0, 9000, False)
        elif r == 2000:
            la = [(-2**31 + 1, -2**31 + 9), (2**31 - 20, 2**31 - 2)]
        elif r == 2001:
            lb = [(-2**31 + 5, -2**31 + 6), (2**31 - 19, 2**31 - 2)]
        A.append(la)
        B.append(lb)
    return _rows_to_oset(A), _rows_to_oset(B)


@pytest.mark.parametrize("schedule", [-1, -5, -6, -7, -8])
@pytest.mark.parametrize("seed", [0, 1])
def test_flat_tile_edge_cases(sx, schedule, seed):
    """Every branch of the flat tile kernel (and the same inputs through the
    row-aware tile kernel), bit-exact against the oracle; the widen and clip
    transforms on load run through the same schedule."""
    from subsetix_amd import _core

    a, b = _flat_edge_pair(seed)
    da, db = to_device(a), to_device(b)
    R = da.row_count
    for op, name in ((0, "union"), (1, "intersection"), (2, "difference")):
        print("flat edges seed", seed, "schedule", schedule, name, flush=True)
        exp = ox.setop(name, a, b)
        out_off, ob, oe, ctx = _core.setop_raw(op, (da.begin, da.end, da.row_offsets),
                                               (db.begin, db.end, db.row_offsets), nrows=R, group=schedule)
        (total,) = _core.read_counts(ctx, ctx.words[1])
        np.testing.assert_array_equal(out_off.cpu().numpy(), exp.off)
        np.testing.assert_array_equal(ob[:total].cpu().numpy(), exp.beg)
        np.testing.assert_array_equal(oe[:total].cpu().numpy(), exp.end)
    # single operand with transforms on load (widen: merge of overlaps; clip:
    # dead intervals), on the rows whose coordinates keep the results in int32
    rows = [list(zip(a.beg[a.off[r]:a.off[r + 1]], a.end[a.off[r]:a.off[r + 1]])) for r in range(R)]
    small = _rows_to_oset([lst if r < 690 else [] for r, lst in enumerate(rows)])
    ds = to_device(small)
    _core.set_group_override(schedule)
    try:
        for hx in (1, 3):
            print("flat edges widen", hx, flush=True)
            assert_set_equal(to_host(sx.dilate_interval_set(ds, hx, 0)), ox.dilate(small, hx, 0))
        print("flat edges clip", flush=True)
        assert_set_equal(to_host(sx.clip_interval_set(ds, width=60, height=R)), ox.clip(small, 60, R))
    finally:
        _core.set_group_override(None)


def _side_by_side(s, t, shift):
    """Rows of s followed, in the same row, by the rows of t moved right by
    shift (twice the intervals per row)."""
    off = s.off.astype(np.int64) + t.off.astype(np.int64)
    beg = np.empty(int(off[-1]), np.int32)
    end = np.empty(int(off[-1]), np.int32)
    for r in range(s.off.size - 1):
        na = s.off[r + 1] - s.off[r]
        o = off[r]
        beg[o:o + na], end[o:o + na] = s.beg[s.off[r]:s.off[r + 1]], s.end[s.off[r]:s.off[r + 1]]
        beg[o + na:off[r + 1]] = t.beg[t.off[r]:t.off[r + 1]] + shift
        end[o + na:off[r + 1]] = t.end[t.off[r]:t.off[r + 1]] + shift
    return ox.make(off, beg, end, s.rows)


@pytest.mark.parametrize("schedule", [None, -7, -8])
def test_uneven_density_stack_vs_oracle(sx, schedule):
    """Stacked instances whose density alternates (every other instance has
    twice the intervals per row), where tiles of equal row counts planned at
    the mean row length overflow in the dense instances: the count-balanced
    tiles (-8), the equal-row tiles (-7) and the default, bit-exact against
    the oracle for the three operations."""
    from subsetix_amd import _core

    rng = np.random.default_rng(11)
    A, B = [], []
    for k in range(10):
        a = random_canonical(rng, 600, 60000, mean_n=30)
        b = random_canonical(rng, 600, 60000, mean_n=30)
        if k % 2:
            a = _side_by_side(a, random_canonical(rng, 600, 60000, mean_n=30), 65536)
            b = _side_by_side(b, random_canonical(rng, 600, 60000, mean_n=30), 65536)
        A.append(a)
        B.append(b)

    def stack(sets):
        offs, begs, ends, base = [], [], [], 0
        for x in sets:
            offs.append(x.off[:-1].astype(np.int64) + base)
            begs.append(x.beg)
            ends.append(x.end)
            base += x.beg.size
        off = np.concatenate(offs + [np.array([base])])
        return ox.make(off, np.concatenate(begs), np.concatenate(ends), np.arange(off.size - 1, dtype=np.int32))

    a, b = stack(A), stack(B)
    da, db = to_device(a), to_device(b)
    for op, name in ((0, "union"), (1, "intersection"), (2, "difference")):
        print("uneven density", schedule, name, flush=True)
        exp = ox.setop(name, a, b)
        out_off, ob, oe, ctx = _core.setop_raw(op, (da.begin, da.end, da.row_offsets),
                                               (db.begin, db.end, db.row_offsets), nrows=da.row_count,
                                               group=schedule)
        (total,) = _core.read_counts(ctx, ctx.words[1])
        np.testing.assert_array_equal(out_off.cpu().numpy(), exp.off)
        np.testing.assert_array_equal(ob[:total].cpu().numpy(), exp.beg)
        np.testing.assert_array_equal(oe[:total].cpu().numpy(), exp.end)


@pytest.mark.parametrize("seed", [3, 4, 5])
def test_default_schedule_heavy_tailed_rows(sx, seed):
    """The default schedule on 3000 identity rows whose lengths are heavy
    tailed (most rows short, a few thousand intervals long, some empty), with
    negative coordinates: tiles of every size, rows that cross the tile
    capacity, bit-exact against the oracle for the three operations."""
    from subsetix_amd import _core

    rng = np.random.default_rng(seed)
    A, B = [], []
    for r in range(3000):
        na = int(min(rng.pareto(1.2) * 4, 12000))
        nb = int(min(rng.pareto(1.2) * 4, 12000))
        base = int(rng.integers(-3000, 3000))
        A.append(_row_runs(rng, base, na, r % 11 == 0))
        B.append(_row_runs(rng, base + int(rng.integers(-50, 50)), nb, r % 13 == 0))
    a, b = _rows_to_oset(A), _rows_to_oset(B)
    da, db = to_device(a), to_device(b)
    for op, name in ((0, "union"), (1, "intersection"), (2, "difference")):
        print("heavy-tailed seed", seed, name, a.n, b.n, flush=True)
        exp = ox.setop(name, a, b)
        out_off, ob, oe, ctx = _core.setop_raw(op, (da.begin, da.end, da.row_offsets),
                                               (db.begin, db.end, db.row_offsets), nrows=da.row_count, group=0)
        (total,) = _core.read_counts(ctx, ctx.words[1])
        np.testing.assert_array_equal(out_off.cpu().numpy(), exp.off)
        np.testing.assert_array_equal(ob[:total].cpu().numpy(), exp.beg)
        np.testing.assert_array_equal(oe[:total].cpu().numpy(), exp.end)
