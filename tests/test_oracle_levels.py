"""Pins the oracle's level transfers, sparse gradient tag, composite step and
three-level instance engine (CPU only).

Each new oracle function is checked two independent ways: against the
already-pinned reference-shaped oracle functions where the two coincide
(full or aligned geometries: prolong_field / restrict_field, the dense
gradient tag, plain upwind), and against per-cell dictionary models on sparse
sets with touching intervals and missing neighbours, where a dropped term, a
wrong sign, a transposed child order or a missed cell fails."""

import numpy as np
import pytest

from oracle import oracle as ox
from tests.golden_util import canonical_rows, random_canonical, runs_of, to_bitmap


def _cells(s, vals):
    co = ox.cell_offsets(s)
    ir = ox.interval_rows(s)
    out = {}
    for i in range(s.n):
        for k in range(int(s.end[i] - s.beg[i])):
            out[(int(ir[i]), int(s.beg[i]) + k)] = vals[co[i] + k]
    return out


def _write_cells(s, vals, cells):
    co = ox.cell_offsets(s)
    ir = ox.interval_rows(s)
    for i in range(s.n):
        for k in range(int(s.end[i] - s.beg[i])):
            key = (int(ir[i]), int(s.beg[i]) + k)
            if key in cells:
                vals[co[i] + k] = cells[key]


def _transfer_model(mode, target, dst, dvals, src, svals, ratio, npdt, accumulate):
    sc = _cells(src, svals)
    dc = _cells(dst, dvals)
    out = {}
    for (y, x) in ox.cells(target):
        if (y, x) not in dc:
            continue
        if mode == 0:
            key = (y // ratio, x // ratio)
            if key in sc:
                out[(y, x)] = sc[key]
            continue
        if npdt == np.float32 and accumulate == "f32":
            acc = np.float32(0)
            for dy in range(ratio):
                for dx in range(ratio):
                    v = sc.get((y * ratio + dy, x * ratio + dx))
                    if v is not None:
                        acc = np.float32(acc + np.float32(v))
            out[(y, x)] = np.float32(acc * np.float32(1.0 / (ratio * ratio)))
        else:
            acc = 0.0
            for dy in range(ratio):
                for dx in range(ratio):
                    v = sc.get((y * ratio + dy, x * ratio + dx))
                    if v is not None:
                        acc += float(v)
            out[(y, x)] = npdt(acc * (1.0 / (ratio * ratio)))
    res = dvals.copy()
    _write_cells(dst, res, out)
    return res


@pytest.mark.parametrize("npdt,acc", [(np.float32, "f32"), (np.float32, "f64"), (np.float64, "f64")])
@pytest.mark.parametrize("touching", [False, True])
def test_transfers_vs_cell_model(npdt, acc, touching):
    rng = np.random.default_rng(7 + int(touching))
    for ratio in (1, 2, 3):
        coarse = random_canonical(rng, 20, 50, mean_n=3, touching=touching)
        fine = random_canonical(rng, 20 * ratio, 50 * ratio, mean_n=5, touching=touching)
        tf = ox.intersection(random_canonical(rng, 20 * ratio, 50 * ratio, mean_n=4), fine)
        tc = random_canonical(rng, 20, 50, mean_n=3)  # cells partly outside coarse
        cv = rng.standard_normal(int(ox.cell_offsets(coarse)[-1])).astype(npdt)
        fv = rng.standard_normal(int(ox.cell_offsets(fine)[-1])).astype(npdt)
        got = fv.copy()
        ox.prolong_onto(tf, fine, got, coarse, cv, ratio)
        np.testing.assert_array_equal(got, _transfer_model(0, tf, fine, fv, coarse, cv, ratio, npdt, acc))
        got = cv.copy()
        ox.restrict_onto(tc, coarse, got, fine, fv, ratio, acc)
        np.testing.assert_array_equal(got, _transfer_model(1, tc, coarse, cv, fine, fv, ratio, npdt, acc))
        assert not np.array_equal(got, cv) or tc.n == 0


@pytest.mark.parametrize("npdt,acc", [(np.float32, "f32"), (np.float32, "f64"), (np.float64, "f64")])
def test_transfers_equal_reference_field_ops_on_aligned_sets(npdt, acc):
    """On aligned geometries the transfers are the reference's prolong_field
    and restrict_field (multilevel.py:365-479) followed by a scatter."""
    rng = np.random.default_rng(3)
    # rows with intervals only: the reference raises on empty coarse rows here
    coarse = ox.compact_rows(random_canonical(rng, 30, 80, mean_n=4))
    fine = ox.prolong_set(coarse, 2)
    cv = rng.random(int(ox.cell_offsets(coarse)[-1])).astype(npdt)
    _, fexp = ox.prolong_field(coarse, cv, 2, dtype=npdt)
    got = np.zeros_like(fexp)
    ox.prolong_onto(fine, fine, got, coarse, cv, 2)
    np.testing.assert_array_equal(got, fexp)
    fv = rng.random(fexp.size).astype(npdt)
    cset, cexp = ox.restrict_field(fine, fv, 2, "mean", acc)
    assert canonical_rows(cset) == canonical_rows(coarse)
    got = np.zeros_like(cv)
    ox.restrict_onto(coarse, coarse, got, fine, fv, 2, acc)
    np.testing.assert_array_equal(got, cexp)


def _grad_tag_model(s, vals, t, npdt):
    cells = _cells(s, vals)
    tagged = set()
    for (y, x), c in cells.items():
        l, r, d, u = (y, x - 1), (y, x + 1), (y - 1, x), (y + 1, x)
        gx = gy = npdt(0)
        if l in cells and r in cells:
            gx = npdt((cells[r] - cells[l]) * npdt(0.5))
        elif r in cells:
            gx = npdt(cells[r] - c)
        elif l in cells:
            gx = npdt(c - cells[l])
        if d in cells and u in cells:
            gy = npdt((cells[u] - cells[d]) * npdt(0.5))
        elif u in cells:
            gy = npdt(cells[u] - c)
        elif d in cells:
            gy = npdt(c - cells[d])
        if np.sqrt(npdt(npdt(gx * gx) + npdt(gy * gy))) >= npdt(t):
            tagged.add((y, x))
    return tagged


@pytest.mark.parametrize("npdt", [np.float32, np.float64])
def test_grad_tag_vs_cell_model(npdt):
    rng = np.random.default_rng(11)
    s = random_canonical(rng, 25, 70, mean_n=4, touching=True)
    vals = rng.random(int(ox.cell_offsets(s)[-1])).astype(npdt)
    for t in (0.05, 0.3, 0.6):
        got = ox.grad_tag(s, vals, t)
        np.testing.assert_array_equal(got.rows, s.rows)
        assert ox.cells(got) == _grad_tag_model(s, vals, t, npdt), t
        # runs are maximal (canonical) per row, also across touching intervals
        bm = to_bitmap(got, 0, 80, 0, 25)
        assert canonical_rows(got) == {y: runs_of(bm[y]) for y in range(25)}


def test_grad_tag_equals_dense_gradient_tag():
    """On a full grid the sparse tag is the reference's dense gradient tag
    (regrid.py:336-350)."""
    rng = np.random.default_rng(5)
    w, h = 37, 23
    d = rng.random((h, w)).astype(np.float32)
    got = ox.grad_tag(ox.full(w, h), d.reshape(-1), 0.3)
    exp = ox.gradient_tag(d, w, h, 0.3)
    assert canonical_rows(got) == canonical_rows(exp)


def _composite_model(c_set, cv, ext, fv, ghost, refine, W, H, vel, dt, npdt):
    cc = _cells(c_set, cv)
    fv = fv.copy()
    _write_cells(ext, fv, {k: cc[(k[0] // 2, k[1] // 2)] for k in ox.cells(ghost)})
    cv2 = ox.upwind(c_set, cv, width=W, height=H, a=vel[0], b=vel[1], dt=dt, dx=1.0 / W, dy=1.0 / H)
    fv2 = ox.upwind(ext, fv, width=2 * W, height=2 * H, a=vel[0], b=vel[1], dt=dt, dx=0.5 / W, dy=0.5 / H)
    fc = _cells(ext, fv2)
    out = {}
    for (y, x) in ox.cells(refine):
        acc = 0.0
        for dy in range(2):
            for dx in range(2):
                acc += float(fc.get((2 * y + dy, 2 * x + dx), 0.0))
        out[(y, x)] = npdt(acc * 0.25)
    _write_cells(c_set, cv2, out)
    return cv2, fv2


@pytest.mark.parametrize("npdt", [np.float32, np.float64])
def test_composite_step_vs_cell_model(npdt):
    W = H = 24
    refine = ox.make([0] * 6 + list(range(1, 11)) + [10] * 9, [5, 4, 4, 5, 6, 6, 7, 8, 9, 9],
                     [12, 13, 14, 14, 15, 16, 16, 15, 14, 13])
    comp = ox.Composite(W, H, refine, velocity=(0.8, -0.6), cfl=0.5, dtype=npdt)
    rng = np.random.default_rng(2)
    comp.set_values(rng.random(W * H).astype(npdt), rng.random(comp.fv.size).astype(npdt))
    hc, hf = comp.cv.copy(), comp.fv.copy()
    for _ in range(4):
        comp.step()
        hc, hf = _composite_model(comp.coarse, hc, comp.ext, hf, comp.ghost, comp.refine, W, H, (0.8, -0.6),
                                  comp.dt, npdt)
    np.testing.assert_array_equal(comp.cv, hc)
    np.testing.assert_array_equal(comp.fv, hf)


def test_instance_one_level_is_plain_upwind():
    n = 40
    rng = np.random.default_rng(9)
    u0 = rng.random(n * n).astype(np.float32)
    inst = ox.Instance3(n, u0, (0.7, -0.4), levels=1, dtype=np.float32)
    inst.run(5)
    u = u0.copy()
    dt = 0.5 / (0.7 * n + 0.4 * n)
    a, b, dtf, dx = (float(np.float32(v)) for v in (0.7, -0.4, dt, 1.0 / n))
    for _ in range(5):
        u = ox.upwind(ox.full(n, n), u, width=n, height=n, a=a, b=b, dt=dtf, dx=dx, dy=dx)
    np.testing.assert_array_equal(inst.vals[0], u)


def _dilate_cells(cells, r=1):
    return {(y + dy, x + dx) for (y, x) in cells for dy in range(-r, r + 1) for dx in range(-r, r + 1)}


def test_instance_three_level_nesting():
    """test_amr_grading.py:110-248 nesting rules on the restated regrid:
    level 2 lies inside level 1, no level-2 cell touches (8 neighbours) a
    level-1-resolution cell outside level 1, the parents of the refined
    cells are refined, and the ghost rings are exactly the one-cell rings
    inside the domain."""
    n = 48
    xs = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xs, xs)
    u0 = np.exp(-((X - 0.45) ** 2 + (Y - 0.5) ** 2) / (2 * 0.08 ** 2)).astype(np.float32)
    inst = ox.Instance3(n, u0.reshape(-1), (0.6, 0.3), threshold=(0.02, 0.02))
    for _ in range(2):
        l1 = ox.cells(inst.sets[1])
        l2 = ox.cells(inst.sets[2])
        assert l1 and l2
        l2_as_l1 = {(y // 2, x // 2) for (y, x) in l2}
        assert l2_as_l1 <= l1
        assert _dilate_cells(l2_as_l1) <= l1  # no 8-neighbour contact with cells outside level 1
        assert {(y // 2, x // 2) for (y, x) in l1} == ox.cells(inst.refines[0])
        for lv in (1, 2):
            w = n * 2 ** lv
            ring = {(y, x) for (y, x) in _dilate_cells(ox.cells(inst.sets[lv])) if 0 <= y < w and 0 <= x < w}
            assert ox.cells(inst.ghosts[lv]) == ring - ox.cells(inst.sets[lv])
        inst.run(8)
    assert inst.step_count == 16 and inst.last_regrid == 8
