"""GPU parity: morphology, multilevel sets and fields, upwind stencil."""

import numpy as np
import pytest

from oracle import oracle as ox
from tests.golden_util import assert_set_equal, canonical_rows, oset, random_canonical
from tests.gpu_util import hostset_to_oset, to_device, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import subsetix_amd

    return subsetix_amd


def test_morphology_golden(sx, golden):
    from subsetix_amd import morphology as m

    for case in golden["morphology"]:
        print("morph", case["id"], flush=True)
        s = to_device(oset(case["input"]))
        if case["op"] == "dilate":
            got = m.dilate_interval_set(s, case["hx"], case["hy"])
        elif case["op"] == "translate":
            got = m.translate_interval_set(s, case["dx"], case["dy"])
        elif case["op"] == "ghost":
            got = m.ghost_zones(s, case["hx"], case["hy"], width=case["width"], height=case["height"])
        else:
            got = m.erode_interval_set(s, case["hx"], case["hy"], width=case["width"], height=case["height"])
        assert_set_equal(to_host(got), oset(case["expect"]))
    with pytest.raises(ValueError):
        m.dilate_interval_set(to_device(ox.make([0, 1], [1], [2])), halo_x=-1)
    with pytest.raises(TypeError):
        m.translate_interval_set(to_device(ox.make([0, 1], [0], [1])), dx=1.5)


@pytest.mark.parametrize("seed", [0, 1])
@pytest.mark.parametrize("hx,hy", [(1, 1), (2, 2), (4, 4), (3, 0), (0, 2), (1, 5)])
def test_dilate_erode_vs_oracle(sx, seed, hx, hy):
    from subsetix_amd import morphology as m

    rng = np.random.default_rng(seed * 10 + hx + 7 * hy)
    rows = np.sort(rng.choice(np.arange(-10, 60), size=40, replace=False)).astype(np.int32)
    s = random_canonical(rng, 40, 300, mean_n=6, rows=rows)
    d = to_device(s)
    print("dilate", seed, hx, hy, flush=True)
    assert_set_equal(to_host(m.dilate_interval_set(d, hx, hy)), ox.dilate(s, hx, hy))
    print("erode", seed, hx, hy, flush=True)
    assert_set_equal(to_host(m.erode_interval_set(d, hx, hy, width=250, height=45)),
                     ox.erode(s, hx, hy, width=250, height=45))
    print("clip/ghost", flush=True)
    assert_set_equal(to_host(m.clip_interval_set(d, width=200, height=30)), ox.clip(s, 200, 30))
    assert_set_equal(to_host(m.ghost_zones(d, hx, hy, width=250, height=45)),
                     ox.ghost_zones(s, hx, hy, width=250, height=45))


@pytest.mark.parametrize("r", [1, 2, 4])
def test_setalgebra_config_morphology_full_size(sx, r):
    """Dilate / erode with a square structuring element of radius r on the
    4096 x 65536 set-algebra inputs, bit-exact."""
    from subsetix_amd import morphology as m
    from subsetix_amd import workloads

    ha, _ = workloads.setalgebra_pair()
    a = hostset_to_oset(ha)
    da = to_device(a)
    print("4096x65536 set dilate r", r, flush=True)
    assert_set_equal(to_host(m.dilate_interval_set(da, r, r)), ox.dilate(a, r, r))
    print("4096x65536 set erode r", r, flush=True)
    assert_set_equal(to_host(m.erode_interval_set(da, r, r, width=65536, height=4096)),
                     ox.erode(a, r, r, width=65536, height=4096))


@pytest.mark.parametrize("r", [1, 4])
def test_dilate_stacked_instances_vs_oracle(sx, r):
    """The batched form of the morphology benchmark: set-algebra instances
    stacked with 16 empty rows between them, dilated in one call, equal the
    oracle's dilation of each instance with its rows shifted (instances of
    4096 rows and a few short ones, so row windows cross every gap)."""
    from subsetix_amd import morphology as m
    from subsetix_amd import workloads

    sets = [workloads.setalgebra_pair(seed_a=1 + 2 * k, seed_b=2 + 2 * k)[0] for k in range(3)]
    sets += [workloads.random_rows_set(n, 65536, seed=40 + n) for n in (5, 1, 37)]
    stacked = workloads.stack_rows_gapped(sets, 16)
    print("stacked dilate r", r, "rows", stacked.rows.size, flush=True)
    got = to_host(m.dilate_interval_set(to_device(hostset_to_oset(stacked)), r, r))
    rows, off, beg, end = [], [0], [], []
    shift = 0
    for h in sets:
        d = ox.dilate(hostset_to_oset(h), r, r)
        rows.append(d.rows.astype(np.int64) + shift)
        off.extend((d.off[1:].astype(np.int64) + off[-1]).tolist())
        beg.append(d.beg)
        end.append(d.end)
        shift += int(h.rows[-1]) + 1 + 16
    exp = ox.make(np.array(off), np.concatenate(beg), np.concatenate(end), rows=np.concatenate(rows))
    assert_set_equal(got, exp)


def test_translate_and_boundaries(sx):
    from subsetix_amd import morphology as m

    base = to_device(ox.make([0, 1], [0], [4]))
    assert canonical_rows(to_host(m.interior_for_direction(base, 1, 0)))[0] == [(0, 3)]
    assert canonical_rows(to_host(m.boundary_for_direction(base, 1, 0)))[0] == [(3, 4)]
    assert canonical_rows(to_host(m.boundary_layer(base)))[0] == [(0, 4)]
    sp = to_device(ox.make([0, 1, 2], [0, 0], [3, 3], rows=[2, 5]))
    got = to_host(m.interior_for_direction(sp, 0, 3))
    assert list(got.rows) == [2] and canonical_rows(got)[2] == [(0, 3)]
    layer = to_host(m.boundary_layer(sp))
    assert list(layer.rows) == [2, 5]


def test_multilevel_golden(sx, golden):
    from subsetix_amd import multilevel as ml

    for case in golden["multilevel"]:
        print("ml", case["id"], flush=True)
        s = to_device(oset(case["input"]))
        if case["op"] == "prolong_then_restrict":
            f = ml.prolong_set(s, case["ratio"])
            assert f.row_count == case["fine_rows"]
            got = ml.restrict_set(f, case["ratio"])
        elif case["op"] == "prolong_set":
            got = ml.prolong_set(s, case["ratio"])
        else:
            got = ml.restrict_set(s, case["ratio"])
        assert_set_equal(to_host(got), oset(case["expect"]))


@pytest.mark.parametrize("ratio", [1, 2, 3, 4])
def test_prolong_restrict_sets_vs_oracle(sx, ratio):
    from subsetix_amd import multilevel as ml

    rng = np.random.default_rng(ratio)
    rows = np.sort(rng.choice(np.arange(-20, 80), size=50, replace=False)).astype(np.int32)
    s = random_canonical(rng, 50, 200, mean_n=5, rows=rows)
    d = to_device(s)
    assert_set_equal(to_host(ml.prolong_set(d, ratio)), ox.prolong_set(s, ratio))
    assert_set_equal(to_host(ml.restrict_set(d, ratio)), ox.restrict_set(s, ratio))
    e = ox.make([0, 0, 0], [], [], rows=[-3, 4])
    assert_set_equal(to_host(ml.prolong_set(to_device(e), ratio)), ox.prolong_set(e, ratio))
    assert_set_equal(to_host(ml.restrict_set(to_device(e), ratio)), ox.restrict_set(e, ratio))


def test_fields_golden(sx, golden):
    import torch

    from subsetix_amd import multilevel as ml
    from subsetix_amd.interval_field import create_interval_field

    for case in golden["fields"]:
        print("fields", case["id"], flush=True)
        s = to_device(oset(case["input"]))
        if case["op"] == "prolong_then_restrict_field":
            cf = create_interval_field(s, 0.0, dtype=torch.float32)
            cf.values.copy_(torch.tensor(case["values"], dtype=torch.float32))
            ff = ml.prolong_field(cf, case["ratio"])
            back = ml.restrict_field(ff, case["ratio"], reducer=case["reducer"])
            np.testing.assert_array_equal(back.values.cpu().numpy(), np.array(case["expect_values"], np.float32))
        elif case["op"] == "restrict_field_pattern":
            cf = create_interval_field(s, 0, dtype=torch.int32)
            ff = ml.prolong_field(cf, case["ratio"])
            pat = torch.tensor(case["pattern"], dtype=torch.int32)
            tiled = pat.repeat(ff.values.numel() // pat.numel() + 1)[: ff.values.numel()]
            ff.values.copy_(tiled.to(ff.values.dtype))
            back = ml.restrict_field(ff, case["ratio"], reducer=case["reducer"])
            assert back.values.dtype == torch.float32
            np.testing.assert_array_equal(back.values.cpu().numpy(), np.array(case["expect_values"], np.float32))
        else:
            ff = create_interval_field(s, 1.0, dtype=torch.float32)
            with pytest.raises(ValueError):
                ml.restrict_field(ff, case["ratio"])


@pytest.mark.parametrize("ratio", [2, 3])
def test_prolong_restrict_fields_vs_oracle(sx, ratio):
    import torch

    from subsetix_amd import multilevel as ml
    from subsetix_amd.interval_field import create_interval_field

    rng = np.random.default_rng(100 + ratio)
    s = random_canonical(rng, 60, 500, mean_n=4, rows=np.arange(5, 65))
    d = to_device(s)
    ncell = int(ox.cell_offsets(s)[-1])
    vals = (rng.random(ncell) * 100).astype(np.float32)
    cf = create_interval_field(d, 0.0, dtype=torch.float32)
    cf.values.copy_(torch.from_numpy(vals))
    ff = ml.prolong_field(cf, ratio)
    f_ref, fv_ref = ox.prolong_field(ox.compact_rows(s), vals, ratio)
    assert_set_equal(to_host(ff.interval_set), f_ref)
    np.testing.assert_array_equal(ff.values.cpu().numpy(), fv_ref)
    # perturb fine values, restrict back with every reducer and accumulation
    fvals = (rng.random(fv_ref.size) * 10).astype(np.float32)
    ff.values.copy_(torch.from_numpy(fvals))
    for red in ("mean", "sum", "min", "max"):
        for acc in ("f32", "f64"):
            got = ml.restrict_field(ff, ratio, reducer=red, accumulate=acc)
            cset, cv = ox.restrict_field(f_ref, fvals, ratio, red, acc)
            assert_set_equal(to_host(got.interval_set), cset)
            np.testing.assert_array_equal(got.values.cpu().numpy(), cv)
    ff64 = ml.prolong_field(cf, ratio, dtype=torch.float64)
    ff64.values.copy_(torch.from_numpy(fvals.astype(np.float64) * 1.0000001))
    got = ml.restrict_field(ff64, ratio)
    _, cv = ox.restrict_field(f_ref, fvals.astype(np.float64) * 1.0000001, ratio, "mean")
    np.testing.assert_array_equal(got.values.cpu().numpy(), cv)


def test_prolong_field_empty_coarse_rows(sx):
    """Coarse rows without intervals (the reference raises there): every fine
    cell takes its parent's value (closed form)."""
    import torch

    from subsetix_amd import multilevel as ml
    from subsetix_amd.interval_field import create_interval_field

    s = ox.make([0, 2, 2, 3], [0, 5, 1], [2, 7, 4])
    cf = create_interval_field(to_device(s), 0.0, dtype=torch.float32)
    vals = np.arange(1, 8, dtype=np.float32)
    cf.values.copy_(torch.from_numpy(vals))
    ff = ml.prolong_field(cf, 2)
    fh = to_host(ff.interval_set)
    assert list(fh.rows) == [0, 1, 4, 5]
    fv = ff.values.cpu().numpy()
    co = ox.cell_offsets(fh)
    ir = ox.interval_rows(fh)
    ccells = {}
    cco = ox.cell_offsets(s)
    cir = ox.interval_rows(s)
    for i in range(s.n):
        for k in range(s.end[i] - s.beg[i]):
            ccells[(cir[i], s.beg[i] + k)] = vals[cco[i] + k]
    for i in range(fh.n):
        for k in range(fh.end[i] - fh.beg[i]):
            x = fh.beg[i] + k
            assert fv[co[i] + k] == ccells[(ir[i] // 2, x // 2)]


def _stencil_setup(golden):
    c = golden["stencil"]
    w, h = c["width"], c["height"]
    x0, x1, y0, y1 = c["box"]
    spans = {str(r): [[x0, x1]] for r in range(y0, y1)}
    from tests.golden_util import spans_set

    s = spans_set(h, spans)
    u = np.concatenate([r * 0.25 + np.arange(x1 - x0, dtype=np.float32) * 0.5 for r in range(y0, y1)]).astype(np.float32)
    return c, s, u


def test_stencil_golden(sx, golden):
    import torch

    from subsetix_amd.interval_field import create_interval_field
    from subsetix_amd.interval_stencil import step_upwind_interval, step_upwind_interval_field
    from tests.test_oracle_golden import _stencil_python

    c, s, u = _stencil_setup(golden)
    w, h = c["width"], c["height"]
    kw = dict(width=w, height=h, a=c["a"], b=c["b"], dt=c["dt"], dx=1.0 / w, dy=1.0 / h)
    field = create_interval_field(to_device(s), 0.0, dtype=torch.float32)
    field.values.copy_(torch.from_numpy(u))
    ref = _stencil_python(s, u, **kw)
    out_f = step_upwind_interval_field(field, **kw)
    buf = torch.empty_like(field.values)
    out_v = step_upwind_interval(field, out=buf, row_ids=field.interval_set.interval_rows(), **kw)
    assert out_v is buf
    np.testing.assert_allclose(out_f.values.cpu().numpy(), ref, rtol=c["tol"], atol=c["tol"])
    np.testing.assert_allclose(buf.cpu().numpy(), ref, rtol=c["tol"], atol=c["tol"])
    np.testing.assert_array_equal(buf.cpu().numpy(), ox.upwind(s, u, **kw))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_stencil_vs_oracle(sx, seed):
    import torch

    from subsetix_amd.interval_field import create_interval_field
    from subsetix_amd.interval_stencil import step_upwind_interval

    rng = np.random.default_rng(seed)
    rows = np.sort(rng.choice(np.arange(0, 90), size=70, replace=False)).astype(np.int32)
    s = random_canonical(rng, 70, 400, mean_n=[2, 8, 40][seed], rows=rows, touching=seed == 1)
    d = to_device(s)
    ncell = int(ox.cell_offsets(s)[-1])
    u = rng.random(ncell).astype(np.float32)
    field = create_interval_field(d, 0.0, dtype=torch.float32)
    field.values.copy_(torch.from_numpy(u))
    for a, b in ((0.7, 0.4), (-0.3, 0.9), (0.5, -0.5), (-1.0, -0.25)):
        kw = dict(width=380, height=90, a=a, b=b, dt=0.01, dx=1.0 / 380, dy=1.0 / 90)
        got = step_upwind_interval(field, **kw).cpu().numpy()
        np.testing.assert_array_equal(got, ox.upwind(s, u, **kw))
        got64 = step_upwind_interval(field, compute_dtype=torch.float64, **kw).cpu().numpy()
        ref64 = ox.upwind(s, u.astype(np.float64), **kw)
        np.testing.assert_array_equal(got64, ref64)
    with pytest.raises(ValueError):
        step_upwind_interval(field, width=380, height=50, a=1.0, b=1.0, dt=0.1, dx=1.0, dy=1.0)


def test_mass(sx):
    import torch

    from subsetix_amd.amr2.mass import total_mass

    rng = np.random.default_rng(5)
    for n in (0, 1, 4095, 4096, 100001):
        u = rng.random(n)
        for dt in (torch.float32, torch.float64):
            t = torch.from_numpy(u).to(dt).cuda()
            got = total_mass(t, 0.25)
            exp = ox.mass(u.astype(np.float32 if dt == torch.float32 else np.float64), 0.25)
            assert abs(got - exp) <= 1e-12 * max(1.0, abs(exp))


def _many_small_intervals(rng, nrows, per_row, touching):
    off, beg, end = [0], [], []
    for _ in range(nrows):
        x = int(rng.integers(-3, 3))
        for _ in range(per_row):
            b = x + (int(rng.integers(0, 3)) if touching else int(rng.integers(1, 3)))
            e = b + int(rng.integers(1, 3))
            beg.append(b)
            end.append(e)
            x = e
        off.append(len(beg))
    return ox.make(off, beg, end)


@pytest.mark.parametrize("schedule", ["tiled", "rows"])
def test_stencil_schedules_vs_oracle(sx, schedule):
    """Both stencil schedules, bit-exact: dense rows spanning several cell
    tiles; rows of ~500 tiny (touching) intervals whose tile neighbourhood
    exceeds the LDS stage (global-memory path); rows with gaps."""
    import torch

    from subsetix_amd import interval_stencil as st
    from subsetix_amd.interval_field import create_interval_field

    rng = np.random.default_rng(11)
    cases = [
        (ox.make(list(range(65)), [0] * 64, [3000] * 64), 3000, 64),
        (_many_small_intervals(rng, 12, 500, True), 1500, 12),
        (_many_small_intervals(rng, 9, 700, False), 2600, 9),
    ]
    rows_gap = np.array([0, 1, 2, 5, 6, 9, 10, 11], np.int32)
    cases.append((ox.make(list(range(0, 9)), [-2, 0, 5, 1, 0, 3, 0, 0], [40, 41, 60, 7, 50, 52, 30, 45],
                          rows=rows_gap), 45, 12))
    st.set_stencil_schedule(schedule)
    try:
        for s, w, h in cases:
            ncell = int(ox.cell_offsets(s)[-1])
            u = rng.random(ncell).astype(np.float32)
            field = create_interval_field(to_device(s), 0.0, dtype=torch.float32)
            field.values.copy_(torch.from_numpy(u))
            for a, b in ((0.7, 0.4), (-0.3, -0.9)):
                kw = dict(width=w, height=h, a=a, b=b, dt=0.01, dx=1.0 / w, dy=1.0 / h)
                print(schedule, "cells", ncell, "a,b", a, b, flush=True)
                got = st.step_upwind_interval(field, **kw).cpu().numpy()
                np.testing.assert_array_equal(got, ox.upwind(s, u, **kw))
                got64 = st.step_upwind_interval(field, compute_dtype=torch.float64, **kw).cpu().numpy()
                np.testing.assert_array_equal(got64, ox.upwind(s, u.astype(np.float64), **kw))
    finally:
        st.set_stencil_schedule("auto")


@pytest.mark.parametrize("hy", [1, 2, 3, 4, 5])
def test_dilate_windows_vs_oracle(sx, hy):
    """Dilation on sparse rows with gaps, empty rows kept in the input,
    negative row ids and touching intervals, hx 0..3, against the oracle:
    through dilate_interval_set and through the tiled dense-row kernel and
    the per-row fused kernel directly for every hy they take."""
    rng = np.random.default_rng(40 + hy)
    rows = np.sort(rng.choice(np.arange(-30, 200), 90, replace=False))
    s = random_canonical(rng, 90, 300, mean_n=3, rows=rows, touching=True)
    # empty rows in the input (zero intervals) must not produce output rows
    off = s.off.copy()
    keep = rng.random(s.rows.size) < 0.8
    beg, end, noff = [], [], [0]
    for r in range(s.rows.size):
        if keep[r]:
            beg += list(s.beg[s.off[r]:s.off[r + 1]])
            end += list(s.end[s.off[r]:s.off[r + 1]])
        noff.append(len(beg))
    s2 = ox.make(noff, beg, end, s.rows)
    from subsetix_amd import _lib
    from subsetix_amd.morphology import _dilate_fused, _dilate_tiled

    lo, hi = int(s2.rows[0]), int(s2.rows[-1])
    for hx in (0, 1, 3):
        print("dilate hy", hy, "hx", hx, flush=True)
        exp = ox.dilate(s2, hx, hy)
        assert_set_equal(to_host(sx.dilate_interval_set(to_device(s2), hx, hy)), exp)
        if hy <= _lib.lib().stx_dilate_max_halo():
            assert_set_equal(to_host(_dilate_fused(to_device(s2), hx, hy)), exp)
            assert_set_equal(to_host(_dilate_tiled(to_device(s2), hx, hy, lo, hi)), exp)
    if hy <= _lib.lib().stx_dilate_max_halo():
        # a row range that is not the set's is reported, not silently cropped
        with pytest.raises(_lib.StxError, match="row ids out of range"):
            _dilate_tiled(to_device(s2), 1, hy, lo, hi - 1)


@pytest.mark.parametrize("r", [1, 2, 4])
def test_morph_uneven_density_vs_oracle(sx, r):
    """Dilation and erosion of a stack whose instances alternate between 20
    and 80 intervals per row (600 rows each, 16 empty rows between them):
    the row tiles are balanced by interval count, so the dense instances get
    shorter tiles; bit-exact against the oracle on the whole stack."""
    from subsetix_amd import workloads

    sets = [workloads.random_rows_set(600, 65536, seed=30 + 8 * r + k, mean_count=80 if k % 2 else 20, mean_len=20,
                                      mean_gap=20) for k in range(8)]
    st = hostset_to_oset(workloads.stack_rows_gapped(sets, 16))
    d = to_device(st)
    print("uneven stack r", r, "intervals", st.beg.size, flush=True)
    assert_set_equal(to_host(sx.dilate_interval_set(d, r, r)), ox.dilate(st, r, r))
    W, H = 65536, int(st.rows[-1]) + 1
    got = to_host(sx.erode_interval_set(d, r, r, width=W, height=H))
    assert_set_equal(got, ox.erode(st, r, r, width=W, height=H))


def _rows_set(rows, per_row):
    off, beg, end = [0], [], []
    for iv in per_row:
        for b, e in iv:
            beg.append(b)
            end.append(e)
        off.append(len(beg))
    return ox.make(off, beg, end, np.asarray(rows, dtype=np.int32))


@pytest.mark.parametrize("hy", [1, 2, 4])
def test_dilate_tiles_edge_cases_vs_oracle(sx, hy):
    """The tiled dense-row dilation where a tile leaves the LDS layout and
    merges its rows one thread per row: a tile whose input exceeds an LDS
    buffer (one row of 6000 intervals), doubling steps that outgrow a
    buffer (rows of unit intervals at row-dependent offsets, so that every
    union of shifted rows adds intervals), x coordinates spanning 2^30 (keys
    would not fit), row ids next to the int32 limits; each against the
    oracle, through dilate_interval_set and the tiled kernel directly."""
    from subsetix_amd.morphology import _dilate_tiled

    for name, rows, per, _, _ in _tile_edge_cases(hy):
        s = _rows_set(rows, per)
        for hx in (0, 2):
            print("tiles", name, "hy", hy, "hx", hx, flush=True)
            exp = ox.dilate(s, hx, hy)
            d = to_device(s)
            assert_set_equal(to_host(sx.dilate_interval_set(d, hx, hy)), exp)
            assert_set_equal(to_host(_dilate_tiled(d, hx, hy, int(s.rows[0]), int(s.rows[-1]))), exp)


@pytest.mark.parametrize("hy", [0, 1, 2, 4])
def test_erode_tiles_edge_cases_vs_oracle(sx, hy):
    """The one-launch tiled erosion on sets with holes: rows of long runs
    with a few holes each, one row with 6000 holes (its complement exceeds
    an LDS buffer), holes at row-dependent offsets (the dilated complement
    outgrows a buffer), a domain that cuts through the set (rows and x
    outside [0, W) x [0, H) present), a 2^28-wide domain (keys would not
    fit: every tile erodes row by row); against the oracle."""
    for name, rows, per, W, H in _erode_edge_cases(hy):
        s = _rows_set(rows, per)
        for hx in (0, 2):
            print("erode tiles", name, "hy", hy, "hx", hx, flush=True)
            got = to_host(sx.erode_interval_set(to_device(s), hx, hy, width=W, height=H))
            assert_set_equal(got, ox.erode(s, hx, hy, width=W, height=H))


def _holes(lo, hi, holes):
    """[lo, hi) minus unit holes at the given sorted positions."""
    out, x = [], lo
    for h in holes:
        if h > x:
            out.append((int(x), int(h)))
        x = max(x, h + 1)
    if hi > x:
        out.append((int(x), int(hi)))
    return out


def _erode_edge_cases(hy):
    """(name, rows, per-row intervals, width, height) for the erosion."""
    rng = np.random.default_rng(70 + hy)
    cases = []
    rows = np.arange(0, 300)
    per = [_holes(0, 20000, np.sort(rng.choice(np.arange(1, 19999, 3), 6000 if r == 150 else 30, replace=False)))
           for r in rows]
    cases.append(("6000 holes", rows, per, 20000, 300))
    rows = np.arange(-40, 460)
    per = [_holes(-50, 1200, [16 * k + 2 * (r % 8) for k in range(70)]) for r in rows]
    cases.append(("row-dependent holes", rows, per, 1100, 400))
    rows = np.arange(-30, 250)
    per = [_holes(-100, 900, np.sort(rng.choice(np.arange(0, 900), 12, replace=False))) for _ in rows]
    cases.append(("domain cut", rows, per, 700, 200))
    rows = np.arange(0, 120)
    per = [_holes(-(1 << 20), (1 << 27) + 9 * r, np.sort(rng.choice(np.arange(0, 1 << 26, 4096), 8, replace=False)))
           for r in rows]
    cases.append(("wide domain", rows, per, 1 << 28, 120))
    cases.append(("empty set", np.zeros(0, dtype=np.int64), [], 50, 10))
    rows = np.arange(-100, -90)
    cases.append(("rows above the domain", rows, [[(0, 40)] for _ in rows], 50, 10))
    return cases


def _tile_edge_cases(hy):
    """(name, rows, per-row intervals, erosion width, erosion height)."""
    rng = np.random.default_rng(90 + hy)
    cases = []
    # one long row among ordinary ones
    rows = np.arange(0, 300)
    per = [[(b, b + 2) for b in range(0, 24000, 4)] if r == 150 else
           [(int(b), int(b) + 3) for b in np.sort(rng.choice(np.arange(0, 4000, 5), 30, replace=False))]
           for r in rows]
    cases.append(("long row", rows, per, 20000, 260))
    # unit intervals at offsets that differ from row to row: unions of
    # shifted rows grow by the row count
    rows = np.arange(-50, 450)
    per = [[(16 * k + 2 * (r % 8), 16 * k + 2 * (r % 8) + 1) for k in range(70)] for r in rows]
    cases.append(("growing unions", rows, per, 1000, 400))
    # coordinates spanning 2^30
    rows = np.arange(0, 200)
    per = [[(-(1 << 29) + 7 * r, -(1 << 29) + 7 * r + 5), (3 * r, 3 * r + 2), ((1 << 29) - 20 - r, (1 << 29) - r)]
           for r in rows]
    cases.append(("wide x", rows, per, 1 << 28, 150))
    # row ids next to either end of the int32 range
    top = (1 << 31) - 1 - hy
    for name, rows in (("int32 low rows", np.arange(-(1 << 31) + hy, -(1 << 31) + hy + 40)),
                       ("int32 high rows", np.arange(top - 60, top + 1))):
        per = [[(int(b), int(b) + 2) for b in np.sort(rng.choice(np.arange(0, 600, 3), 12, replace=False))]
               for _ in rows]
        cases.append((name, rows, per, 500, 60))
    return cases
