"""Pins the CPU oracle: the reference's own golden vectors (transcribed into
tests/golden/reference_vectors.json) plus closed forms and brute force on
bitmaps, chosen so a dropped term, a wrong sign or a swapped operand fails."""

import numpy as np
import pytest

from oracle import oracle as ox
from tests.golden_util import (
    assert_set_equal,
    canonical_rows,
    cell_model,
    eval_spec,
    oset,
    random_canonical,
    runs_of,
    spans_set,
    to_bitmap,
)

OPS = {
    "union": ox.union,
    "intersection": ox.intersection,
    "difference": ox.difference,
    "symmetric_difference": ox.symmetric_difference,
}


def test_expressions_golden(golden):
    surf = {k: oset(v) for k, v in golden["surfaces"].items()}
    for case in golden["expressions"]:
        got = eval_spec(case["spec"], surf, OPS)
        assert_set_equal(got, oset(case["expect"]))


def test_expressions_cell_model(golden):
    surf = {k: oset(v) for k, v in golden["surfaces"].items()}
    cells = {k: ox.cells(v) for k, v in surf.items()}
    for group in ("dense", "sparse"):
        for spec in golden["cell_model_specs"][group]:
            got = eval_spec(spec, surf, OPS)
            assert ox.cells(got) == cell_model(spec, cells), spec


def test_morphology_golden(golden):
    for case in golden["morphology"]:
        s = oset(case["input"])
        if case["op"] == "dilate":
            got = ox.dilate(s, case["hx"], case["hy"])
        elif case["op"] == "translate":
            got = ox.translate(s, case["dx"], case["dy"])
        elif case["op"] == "ghost":
            got = ox.ghost_zones(s, case["hx"], case["hy"], width=case["width"], height=case["height"])
        elif case["op"] == "erode":
            got = ox.erode(s, case["hx"], case["hy"], width=case["width"], height=case["height"])
        assert_set_equal(got, oset(case["expect"]))


def test_dilate_superset(golden):
    c = golden["dilate_superset"]
    s = oset(c["input"])
    d = ox.dilate(s, c["hx"], c["hy"])
    dr = canonical_rows(d)
    for row, ivs in canonical_rows(s).items():
        for b, e in ivs:
            assert any(b >= db and e <= de for db, de in dr[row])


def test_multilevel_golden(golden):
    for case in golden["multilevel"]:
        s = oset(case["input"])
        if case["op"] == "prolong_then_restrict":
            f = ox.prolong_set(s, case["ratio"])
            assert f.nrows == case["fine_rows"]
            got = ox.restrict_set(f, case["ratio"])
        elif case["op"] == "prolong_set":
            got = ox.prolong_set(s, case["ratio"])
        else:
            got = ox.restrict_set(s, case["ratio"])
        assert_set_equal(got, oset(case["expect"]))


def test_fields_golden(golden):
    for case in golden["fields"]:
        s = oset(case["input"])
        if case["op"] == "prolong_then_restrict_field":
            f, fv = ox.prolong_field(s, np.array(case["values"], np.float32), case["ratio"])
            _, cv = ox.restrict_field(f, fv, case["ratio"], case["reducer"])
            np.testing.assert_array_equal(cv, np.array(case["expect_values"], np.float32))
        elif case["op"] == "restrict_field_pattern":
            f, fv = ox.prolong_field(s, np.zeros(2, np.int32), case["ratio"])
            pat = np.array(case["pattern"], np.float32)
            fv = np.tile(pat, fv.size // pat.size + 1)[: fv.size]
            _, cv = ox.restrict_field(f, fv, case["ratio"], case["reducer"])
            np.testing.assert_array_equal(cv, np.array(case["expect_values"], np.float32))
        else:
            with pytest.raises(ox.OracleError):
                ox.restrict_field(s, np.ones(int(ox.cell_offsets(s)[-1]), np.float32), case["ratio"])


def _dense_from_case(case):
    w, h = case["width"], case["height"]
    d = np.zeros((h, w), np.float32)
    if "dense" in case:
        d = np.array(case["dense"], np.float32).reshape(h, w)
    if "fill" in case:
        d[:] = case["fill"]
    for y, x, v in case.get("points", []):
        d[y, x] = v
    return d


def tag_fraction(d, w, h, frac_high, epsilon=1e-8):
    """regrid.py:353-383 (value tag of the top frac_high of the positive values)."""
    frac_high = max(0.0, min(1.0, frac_high))
    v = d.reshape(-1).astype(np.float32)
    pos = v[v > np.float32(epsilon)]
    if frac_high <= 0 or pos.size == 0:
        return spans_set(h, {})
    count = max(1, min(pos.size, int(np.ceil(pos.size * frac_high))))
    thr = float(np.partition(pos, pos.size - count)[pos.size - count])
    return ox.threshold(d, w, h, max(thr, epsilon))


def test_regrid_golden(golden):
    for case in golden["regrid"]:
        if case["op"] == "grading":
            s = spans_set(case["height"], case["spans"])
            got = ox.grading(s, padding=case["padding"], mode=case["mode"], width=case["width"], height=case["height"])
        else:
            d = _dense_from_case(case)
            if case["op"] == "tag_value":
                got = ox.threshold(d, case["width"], case["height"], max(case["threshold"], 1e-8))
            else:
                got = tag_fraction(d, case["width"], case["height"], case["frac_high"], case.get("epsilon", 1e-8))
        assert_set_equal(got, spans_set(case["height"], case["expect_spans"]))
    with pytest.raises(ox.OracleError):
        ox.grading(spans_set(4, {}), padding=1, mode="diag", width=4, height=4)


def test_geometry_golden(golden):
    for case in golden["geometry"]:
        w, h, r = case["width"], case["height"], case["ratio"]
        refine = spans_set(h, case["refine_spans"])
        coarse = ox.full(w, h)
        aligned = ox.intersection(refine, coarse)  # fields.py:536
        co = ox.difference(coarse, aligned)        # geometry.py:58-59
        assert_set_equal(co, spans_set(h, case["expect_coarse_only_spans"]))
        fine = ox.prolong_set(aligned, r)          # fields.py:552-555
        if case["expect_fine"] is not None:
            assert_set_equal(fine, oset(case["expect_fine"]))
        else:
            assert fine.n == 0


def test_sync_golden(golden):
    for case in golden["sync"]:
        h, w = case["coarse_hw"]
        coarse = ox.full(w, h)
        fine = ox.full(2 * w, 2 * h)
        cv = np.full(w * h, case["coarse_fill"], np.float32)
        if case.get("fine_arange"):
            fv = np.arange(4 * w * h, dtype=np.float32)
        else:
            fv = np.full(4 * w * h, case["fine_fill"], np.float32)
        refine = ox.intersection(spans_set(h, case["refine_spans"]), coarse)
        ncv, nfv = ox.synchronize(coarse, cv, fine, fv, refine, ratio=2, fill_fine_outside=case["fill_fine_outside"])
        if "expect_coarse_values" in case:
            np.testing.assert_allclose(ncv, np.array(case["expect_coarse_values"], np.float32))
            np.testing.assert_array_equal(nfv, fv)
        else:
            exp_c = np.zeros((h, w), np.float32)
            for y, x in case["expect_coarse_cells_value_2"]:
                exp_c[y, x] = 2.0
            np.testing.assert_array_equal(ncv.reshape(h, w), exp_c)
            y0, y1, x0, x1 = case["expect_fine_box_value_2"]
            exp_f = np.zeros((2 * h, 2 * w), np.float32)
            exp_f[y0:y1, x0:x1] = 2.0
            np.testing.assert_array_equal(nfv.reshape(2 * h, 2 * w), exp_f)


def test_gather_golden(golden):
    for case in golden["gather"]:
        h, w = case["super_hw"]
        sup = ox.full(w, h)
        vals = np.arange(w * h, dtype=np.float32)
        sub = spans_set(h, case["subset_spans"])
        np.testing.assert_array_equal(ox.gather(sup, vals, sub), np.array(case["expect_values"], np.float32))


def _stencil_python(s, u, width, height, a, b, dt, dx, dy):
    """Per-cell Python upwind of test_interval_stencil.py:15-61 (double math)."""
    cells = {}
    co = ox.cell_offsets(s)
    ir = ox.interval_rows(s)
    for i in range(s.n):
        for k in range(int(s.end[i] - s.beg[i])):
            cells[(int(ir[i]), int(s.beg[i]) + k)] = float(u[co[i] + k])
    out = np.empty_like(u)
    for i in range(s.n):
        row = int(ir[i])
        for k in range(int(s.end[i] - s.beg[i])):
            col = int(s.beg[i]) + k
            c = cells[(row, col)]
            left = cells.get((row, col - 1), 0.0)
            right = cells.get((row, col + 1), 0.0)
            down = cells.get((row - 1, col), 0.0)
            up = cells.get((row + 1, col), 0.0)
            dudx = (c - left) / dx if a >= 0 else (right - c) / dx
            dudy = (c - down) / dy if b >= 0 else (up - c) / dy
            out[co[i] + k] = c - dt * (a * dudx + b * dudy)
    return out


def stencil_case(golden):
    c = golden["stencil"]
    w, h = c["width"], c["height"]
    x0, x1, y0, y1 = c["box"]
    spans = {str(r): [[x0, x1]] for r in range(y0, y1)}
    s = spans_set(h, spans)
    u = np.concatenate([r * 0.25 + np.arange(x1 - x0, dtype=np.float32) * 0.5 for r in range(y0, y1)]).astype(np.float32)
    return c, s, u


def test_stencil_golden(golden):
    c, s, u = stencil_case(golden)
    w, h = c["width"], c["height"]
    kw = dict(width=w, height=h, a=c["a"], b=c["b"], dt=c["dt"], dx=1.0 / w, dy=1.0 / h)
    ref = _stencil_python(s, u, **kw)
    got32 = ox.upwind(s, u, **kw)
    np.testing.assert_allclose(got32, ref, rtol=c["tol"], atol=c["tol"])
    got64 = ox.upwind(s, u.astype(np.float64), **kw)
    np.testing.assert_allclose(got64, ref.astype(np.float64), rtol=1e-6, atol=1e-6)


# ----------------------------------------------------------------------------
# Closed forms / brute force.

@pytest.mark.parametrize("seed", range(6))
def test_setops_bitmap(seed):
    rng = np.random.default_rng(seed)
    a = random_canonical(rng, 24, 120)
    b = random_canonical(rng, 24, 120)
    A, B = to_bitmap(a, 0, 130, 0, 24), to_bitmap(b, 0, 130, 0, 24)
    for op, model in (("union", A | B), ("intersection", A & B), ("difference", A & ~B)):
        got = canonical_rows(ox.setop(op, a, b))
        for y in range(24):
            assert got[y] == runs_of(model[y]), (op, y)


def test_setops_touching_splits():
    # Intersection/difference keep the splits of touching inputs
    # (kernels.py:67-113): pieces per overlapping pair / per lhs interval.
    a = ox.make([0, 2], [0, 2], [2, 4])
    b = ox.make([0, 1], [1], [3])
    assert canonical_rows(ox.intersection(a, b))[0] == [(1, 2), (2, 3)]
    assert canonical_rows(ox.difference(a, b))[0] == [(0, 1), (3, 4)]
    assert canonical_rows(ox.union(a, b))[0] == [(0, 4)]
    c = ox.make([0, 2], [0, 3], [3, 6])
    assert canonical_rows(ox.difference(ox.make([0, 1], [0], [6]), c))[0] == []
    assert canonical_rows(ox.intersection(ox.make([0, 1], [0], [6]), c))[0] == [(0, 3), (3, 6)]


def _dilate_bitmap(bm, hx, hy):
    out = np.zeros_like(bm)
    H, W = bm.shape
    for dy in range(-hy, hy + 1):
        for dx in range(-hx, hx + 1):
            src = bm[max(0, -dy):H - max(0, dy), max(0, -dx):W - max(0, dx)]
            out[max(0, dy):max(0, dy) + src.shape[0], max(0, dx):max(0, dx) + src.shape[1]] |= src
    return out


@pytest.mark.parametrize("seed", range(4))
@pytest.mark.parametrize("hx,hy", [(1, 1), (2, 0), (0, 2), (3, 1)])
def test_dilate_erode_bitmap(seed, hx, hy):
    rng = np.random.default_rng(100 + seed)
    W, H, P = 80, 20, 8
    s = random_canonical(rng, H, W - 4, mean_n=3)
    bm = to_bitmap(s, -P, W + P, -P, H + P)
    d = canonical_rows(ox.dilate(s, hx, hy))
    dm = _dilate_bitmap(bm, hx, hy)
    for yy in range(-P, H + P):
        assert d.get(yy, []) == [(b - P, e - P) for b, e in runs_of(dm[yy + P])], yy
    # erode: p in result iff no q in domain \ S within the box (exterior = inside)
    e = canonical_rows(ox.erode(s, hx, hy, width=W, height=H))
    sb = to_bitmap(s, 0, W, 0, H)
    holes = _dilate_bitmap(~sb, hx, hy)
    for y in range(H):
        assert e[y] == runs_of(~holes[y]), y


def test_clip_and_prolong_restrict_closed_forms():
    rng = np.random.default_rng(7)
    s = random_canonical(rng, 30, 60, rows=np.arange(-5, 25))
    c = ox.clip(s, 40, 20)
    np.testing.assert_array_equal(c.rows, np.arange(20))
    bm = to_bitmap(s, 0, 40, 0, 20)
    cr = canonical_rows(c)
    for y in range(20):
        assert cr[y] == runs_of(bm[y])
    # prolong: [bR, eR) on rows rR..rR+R-1; restrict: floor/ceil division
    for R in (2, 3, 4):
        f = ox.prolong_set(s, R)
        fr = canonical_rows(f)
        for row, ivs in canonical_rows(s).items():
            for k in range(R):
                assert fr.get(row * R + k, []) == [(b * R, e * R) for b, e in ivs] or not ivs
        back = ox.restrict_set(f, R)
        nonempty = {r: v for r, v in canonical_rows(s).items() if v}
        assert canonical_rows(back) == nonempty
    t = ox.make([0, 1, 2], [-3, 5], [-1, 6], rows=[-3, -2])
    r2 = ox.restrict_set(t, 2)
    assert canonical_rows(r2) == {-2: [(-2, 0)], -1: [(2, 3)]}


def test_restrict_field_accumulations():
    rng = np.random.default_rng(3)
    c = ox.make([0, 1, 2], [0, 3], [5, 9])
    f = ox.prolong_set(c, 2)
    fv = rng.random(int(ox.cell_offsets(f)[-1])).astype(np.float32) * 1000
    _, m32 = ox.restrict_field(f, fv, 2, "mean", "f32")
    _, m64 = ox.restrict_field(f, fv, 2, "mean", "f64")
    _, md = ox.restrict_field(f, fv.astype(np.float64), 2, "mean")
    # closed form: mean over the 2x2 children, children order dy-major
    co, fo = ox.cell_offsets(c), ox.cell_offsets(f)
    exp32, exp64 = [], []
    for i in range(c.n):
        for k in range(int(c.end[i] - c.beg[i])):
            ch = [fv[fo[2 * i + dy] + 2 * k + dx] for dy in range(2) for dx in range(2)]
            acc = np.float32(0)
            for v in ch:
                acc = np.float32(acc + v)
            exp32.append(np.float32(acc * np.float32(0.25)))
            exp64.append(sum(float(v) for v in ch) * 0.25)
    np.testing.assert_array_equal(m32, np.array(exp32, np.float32))
    np.testing.assert_array_equal(m64, np.array(exp64).astype(np.float32))
    np.testing.assert_array_equal(md, np.array(exp64))


def test_threshold_and_gradient_closed_form():
    rng = np.random.default_rng(11)
    for w, h in ((1, 5), (33, 7), (64, 3), (65, 4), (130, 2)):
        d = rng.random((h, w)).astype(np.float32)
        d[d < 0.2] = np.nan
        s = ox.threshold(d, w, h, 0.5)
        got = canonical_rows(s)
        for y in range(h):
            assert got[y] == runs_of(d[y] >= np.float32(0.5))
        g = ox.gradient(d, w, h).reshape(h, w)
        gx = np.zeros_like(d)
        gy = np.zeros_like(d)
        if w > 1:
            gx[:, 1:-1] = (d[:, 2:] - d[:, :-2]) * np.float32(0.5)
            gx[:, 0] = d[:, 1] - d[:, 0]
            gx[:, -1] = d[:, -1] - d[:, -2]
        if h > 1:
            gy[1:-1] = (d[2:] - d[:-2]) * np.float32(0.5)
            gy[0] = d[1] - d[0]
            gy[-1] = d[-1] - d[-2]
        np.testing.assert_array_equal(g, np.sqrt(gx * gx + gy * gy))


def test_upwind_f64_closed_form():
    # Two separated rows, one interval each: zero exterior gives closed form.
    s = ox.make([0, 1, 2], [2, 2], [5, 5], rows=[1, 3])
    u = np.array([1.0, 2.0, 4.0, 8.0, 16.0, 32.0], np.float64)
    out = ox.upwind(s, u, width=10, height=6, a=1.0, b=1.0, dt=0.1, dx=1.0, dy=1.0)
    exp = []
    for r in range(2):
        row = u[3 * r:3 * r + 3]
        for k in range(3):
            left = row[k - 1] if k > 0 else 0.0
            exp.append(row[k] - 0.1 * (1.0 * (row[k] - left) + 1.0 * (row[k] - 0.0)))
    np.testing.assert_allclose(out, exp, rtol=0, atol=1e-15)


def test_mass_closed_form():
    """FP64 mass: sum(u) * area against an exactly rounded sum (math.fsum)
    and against integer-valued data whose sum is exact."""
    import math

    rng = np.random.default_rng(9)
    for n in (0, 1, 4095, 4096, 4097, 100003):
        u = rng.random(n)
        got = ox.mass(u, 0.25)
        assert abs(got - math.fsum(u) * 0.25) <= 1e-12 * max(1.0, math.fsum(u))
        u32 = u.astype(np.float32)
        assert abs(ox.mass(u32, 0.5) - math.fsum(u32.astype(np.float64)) * 0.5) <= 1e-12 * max(1.0, n)
    k = np.arange(1000, dtype=np.float64)
    assert ox.mass(k, 2.0) == 2.0 * 999 * 1000 / 2
