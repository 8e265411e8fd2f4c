"""GPU parity: fused level transfers, sparse gradient tagging, the sparse
two-level composite step (closed forms / oracle restatements)."""

import numpy as np
import pytest

from oracle import oracle as ox
from tests.golden_util import assert_set_equal, canonical_rows, random_canonical, runs_of
from tests.gpu_util import to_device, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import subsetix_amd

    return subsetix_amd


def _cells(s, vals):
    co = ox.cell_offsets(s)
    ir = ox.interval_rows(s)
    d = {}
    for i in range(s.n):
        for k in range(int(s.end[i] - s.beg[i])):
            d[(int(ir[i]), int(s.beg[i]) + k)] = vals[co[i] + k]
    return d


def _field(sx, s, vals, dtype):
    import torch

    f = sx.create_interval_field(to_device(s), 0.0, dtype=dtype)
    f.values.copy_(torch.from_numpy(np.asarray(vals)).to(dtype))
    return f


@pytest.mark.parametrize("schedule", ["segments", "rows"])
@pytest.mark.parametrize("dt", ["f32", "f64"])
@pytest.mark.parametrize("touching", [False, True])
def test_prolong_restrict_into(sx, dt, schedule, touching):
    """Both transfer schedules against a per-cell model: unaligned fine sets
    (children partly missing), touching intervals, sparse targets."""
    import torch

    from subsetix_amd import levels
    from subsetix_amd.levels import prolong_into, restrict_into

    levels.set_transfer_schedule(schedule)
    try:
        _prolong_restrict_case(sx, dt, touching)
    finally:
        levels.set_transfer_schedule("auto")


def _prolong_restrict_case(sx, dt, touching):
    import torch

    from subsetix_amd.levels import prolong_into, restrict_into

    print("transfer case", dt, touching, flush=True)
    tdt = torch.float32 if dt == "f32" else torch.float64
    npdt = np.float32 if dt == "f32" else np.float64
    rng = np.random.default_rng(1 + int(touching))
    coarse = random_canonical(rng, 30, 60, mean_n=3, rows=np.arange(30), touching=touching)
    fine = random_canonical(rng, 60, 120, mean_n=5, rows=np.arange(60), touching=touching)
    target_f = random_canonical(rng, 60, 120, mean_n=4, rows=np.arange(60))
    target_f = ox.intersection(target_f, fine)
    cv = rng.random(int(ox.cell_offsets(coarse)[-1])).astype(npdt)
    fv = rng.random(int(ox.cell_offsets(fine)[-1])).astype(npdt)
    cf, ff = _field(sx, coarse, cv, tdt), _field(sx, fine, fv, tdt)
    prolong_into(to_device(target_f), ff, cf, 2)
    cc = _cells(coarse, cv)
    exp = fv.copy()
    co = ox.cell_offsets(fine)
    ir = ox.interval_rows(fine)
    tcells = set(ox.cells(target_f))
    for i in range(fine.n):
        for k in range(int(fine.end[i] - fine.beg[i])):
            y, x = int(ir[i]), int(fine.beg[i]) + k
            if (y, x) in tcells and (y // 2, x // 2) in cc:
                exp[co[i] + k] = cc[(y // 2, x // 2)]
    np.testing.assert_array_equal(ff.values.cpu().numpy(), exp)
    # ratio-1 prolong (a copy): target cells outside the source keep their value
    other = random_canonical(rng, 60, 120, mean_n=5, rows=np.arange(60), touching=not touching)
    ov = rng.random(int(ox.cell_offsets(other)[-1])).astype(npdt)
    of = _field(sx, other, ov, tdt)
    target_1 = ox.intersection(random_canonical(rng, 60, 120, mean_n=6, rows=np.arange(60)), fine)
    prolong_into(to_device(target_1), ff, of, 1)
    oc = _cells(other, ov)
    t1 = set(ox.cells(target_1))
    for i in range(fine.n):
        for k in range(int(fine.end[i] - fine.beg[i])):
            y, x = int(ir[i]), int(fine.beg[i]) + k
            if (y, x) in t1 and (y, x) in oc:
                exp[co[i] + k] = oc[(y, x)]
    np.testing.assert_array_equal(ff.values.cpu().numpy(), exp)
    # restrict onto a coarse target
    target_c = ox.intersection(random_canonical(rng, 30, 60, mean_n=3, rows=np.arange(30)), coarse)
    fv2 = rng.random(fv.size).astype(npdt)
    ff.values.copy_(torch.from_numpy(fv2))
    cf2 = _field(sx, coarse, cv, tdt)
    restrict_into(to_device(target_c), cf2, ff, 2, accumulate="f64")
    fc = _cells(fine, fv2)
    expc = cv.copy()
    cco = ox.cell_offsets(coarse)
    cir = ox.interval_rows(coarse)
    tc = set(ox.cells(target_c))
    for i in range(coarse.n):
        for k in range(int(coarse.end[i] - coarse.beg[i])):
            y, x = int(cir[i]), int(coarse.beg[i]) + k
            if (y, x) in tc:
                acc = 0.0
                for dy in range(2):
                    for dx in range(2):
                        acc += float(fc.get((2 * y + dy, 2 * x + dx), 0.0))
                expc[cco[i] + k] = npdt(acc * 0.25)
    np.testing.assert_array_equal(cf2.values.cpu().numpy(), expc)


def _grad_tag_model(s, vals, t):
    cells = _cells(s, vals)
    tagged = set()
    for (y, x), c in cells.items():
        l, r, d, u = (y, x - 1), (y, x + 1), (y - 1, x), (y + 1, x)
        gx = gy = 0.0
        if l in cells and r in cells:
            gx = (cells[r] - cells[l]) * 0.5
        elif r in cells:
            gx = cells[r] - c
        elif l in cells:
            gx = c - cells[l]
        if d in cells and u in cells:
            gy = (cells[u] - cells[d]) * 0.5
        elif u in cells:
            gy = cells[u] - c
        elif d in cells:
            gy = c - cells[d]
        gx, gy = np.float32(gx), np.float32(gy)
        if np.sqrt(np.float32(gx * gx) + np.float32(gy * gy)) >= np.float32(t):
            tagged.add((y, x))
    return tagged


@pytest.mark.parametrize("schedule", ["rows", "seg"])
def test_gradient_tag_sparse_and_dense(sx, schedule):
    """Per-cell model (sparse set, touching intervals) and the dense oracle;
    'seg' runs the segment-table kernel (bit masks), 'rows' the row-group
    kernel."""
    import torch

    from subsetix_amd.interval_stencil import build_stencil_plan
    from subsetix_amd.levels import gradient_tag

    def tag(s, vals, t, width):
        f = _field(sx, s, vals, torch.float32)
        plan = build_stencil_plan(f.interval_set, f.interval_cell_offsets, width) if schedule == "seg" else None
        return to_host(gradient_tag(f, t, plan=plan))

    print("grad tag", schedule, flush=True)
    rng = np.random.default_rng(9)
    s = random_canonical(rng, 40, 80, mean_n=4, rows=np.arange(40), touching=True)
    v = rng.random(int(ox.cell_offsets(s)[-1])).astype(np.float32)
    got = tag(s, v, 0.3, 80)
    assert ox.cells(got) == _grad_tag_model(s, v, 0.3)
    # runs are maximal within adjacency
    for row, ivs in canonical_rows(got).items():
        for (b0, e0), (b1, e1) in zip(ivs, ivs[1:]):
            assert e0 < b1
    # dense: identical to the dense gradient tag of the reference formula
    w, h = 70, 33
    d = rng.random((h, w)).astype(np.float32)
    dense = ox.full(w, h)
    got = tag(dense, d.reshape(-1), 0.4, w)
    assert_set_equal(got, ox.gradient_tag(d, w, h, 0.4))


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_gradient_tag_seg_equals_rows(sx, dt):
    """Segment kernel == row-group kernel, bit for bit: several cell tiles,
    sparse rows with gaps, touching intervals, long tagged runs crossing
    32-cell words and tiles, an empty field."""
    import torch

    from subsetix_amd.interval_stencil import build_stencil_plan
    from subsetix_amd.levels import gradient_tag

    tdt = torch.float32 if dt == "f32" else torch.float64
    npdt = np.float32 if dt == "f32" else np.float64
    rng = np.random.default_rng(23)
    rows = np.sort(rng.choice(400, 300, replace=False))
    cases = [(random_canonical(rng, 300, 700, mean_n=12, rows=rows, touching=True), 700),
             (ox.full(3000, 7), 3000)]
    for s, w in cases:
        n = int(ox.cell_offsets(s)[-1])
        x = np.linspace(0.0, 40.0, n)
        vals = (np.sin(x) * (rng.random(n) < 0.8) + 0.01 * rng.random(n)).astype(npdt)
        f = _field(sx, s, vals, tdt)
        plan = build_stencil_plan(f.interval_set, f.interval_cell_offsets, w)
        for t in (0.02, 0.3):
            print("grad tag seg vs rows", dt, n, t, flush=True)
            a = to_host(gradient_tag(f, t, plan=plan))
            b = to_host(gradient_tag(f, t))
            for name in ("rows", "off", "beg", "end"):
                np.testing.assert_array_equal(getattr(a, name), getattr(b, name), err_msg=name)
            assert a.n > 0


def test_gradient_tag_seg_rejects_cells_outside_domain(sx):
    import torch

    from subsetix_amd._lib import StxError
    from subsetix_amd.interval_stencil import build_stencil_plan
    from subsetix_amd.levels import gradient_tag

    s = ox.make([0, 1, 2], [0, 5], [40, 60])
    f = _field(sx, s, np.ones(95, np.float32), torch.float32)
    plan = build_stencil_plan(f.interval_set, f.interval_cell_offsets, 50)  # row 1 reaches x = 59
    with pytest.raises(StxError, match="domain"):
        gradient_tag(f, 0.1, plan=plan)


def _oracle_composite_step(c_set, cv, ext, fv, ghost, refine, W, H, vel, dt, dtype):
    """prolong ghosts, upwind both levels, restrict (FP64 sums) - on dicts."""
    npdt = dtype
    cc = _cells(c_set, cv)
    fco = ox.cell_offsets(ext)
    fir = ox.interval_rows(ext)
    gcells = ox.cells(ghost)
    fv = fv.copy()
    for i in range(ext.n):
        for k in range(int(ext.end[i] - ext.beg[i])):
            y, x = int(fir[i]), int(ext.beg[i]) + k
            if (y, x) in gcells:
                fv[fco[i] + k] = cc[(y // 2, x // 2)]
    cv2 = ox.upwind(c_set, cv, width=W, height=H, a=vel[0], b=vel[1], dt=dt, dx=1.0 / W, dy=1.0 / H)
    fv2 = ox.upwind(ext, fv, width=2 * W, height=2 * H, a=vel[0], b=vel[1], dt=dt, dx=0.5 / W, dy=0.5 / H)
    fc = _cells(ext, fv2)
    cco = ox.cell_offsets(c_set)
    cir = ox.interval_rows(c_set)
    rc = ox.cells(refine)
    for i in range(c_set.n):
        for k in range(int(c_set.end[i] - c_set.beg[i])):
            y, x = int(cir[i]), int(c_set.beg[i]) + k
            if (y, x) in rc:
                acc = 0.0
                for dy in range(2):
                    for dx in range(2):
                        acc += float(fc.get((2 * y + dy, 2 * x + dx), 0.0))
                cv2[cco[i] + k] = npdt(acc * 0.25)
    return cv2, fv2


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_composite_steps_vs_oracle(sx, dt):
    import torch

    from subsetix_amd.amr2.composite import CompositeTwoLevel

    tdt = torch.float32 if dt == "f32" else torch.float64
    npdt = np.float32 if dt == "f32" else np.float64
    W = H = 24
    refine = ox.make([0] * 6 + list(range(1, 11)) + [10] * 9, [5, 4, 4, 5, 6, 6, 7, 8, 9, 9],
                     [12, 13, 14, 14, 15, 16, 16, 15, 14, 13])
    comp = CompositeTwoLevel(W, H, to_device(refine), dtype=tdt, velocity=(0.8, -0.6), cfl=0.5)
    rng = np.random.default_rng(2)
    cv = rng.random(W * H).astype(npdt)
    fv = rng.random(comp.fine_field.values.numel()).astype(npdt)
    comp.set_values(torch.from_numpy(cv), torch.from_numpy(fv))
    c_set, ext, ghost, ref = ox.full(W, H), to_host(comp.ext), to_host(comp.ghost), to_host(comp.refine)
    hc = comp.coarse.values.cpu().numpy()
    hf = comp.fine_field.values.cpu().numpy()
    for _ in range(5):
        comp.step()
        hc, hf = _oracle_composite_step(c_set, hc, ext, hf, ghost, ref, W, H, (0.8, -0.6), comp.dt, npdt)
    np.testing.assert_array_equal(comp.coarse.values.cpu().numpy(), hc)
    np.testing.assert_array_equal(comp.fine_field.values.cpu().numpy(), hf)
    assert abs(comp.mass() - ox.mass(hc, 1.0 / W / H)) < 1e-12


def _ulp_neighbours(t, npdt, k=3):
    """t and its k nearest representable neighbours on each side."""
    out = [npdt(t)]
    lo = hi = npdt(t)
    for _ in range(k):
        lo = np.nextafter(lo, npdt(-np.inf))
        hi = np.nextafter(hi, npdt(np.inf))
        out += [lo, hi]
    return out


@pytest.mark.parametrize("dt", ["f32", "f64"])
def test_gradient_tag_threshold_boundaries(sx, dt):
    """|grad u| >= t at the rounding boundary: two-cell rows [0, g) with no
    vertical neighbours give gx = g exactly and gx^2 + gy^2 = fl(g^2), so g
    within a few ulps of t puts the sum on both sides of the smallest s with
    sqrt(s) >= t; plus thresholds 0, -1, 1e-30, inf and nan.  Segment table,
    row groups and the oracle (sqrt(s) >= t, correctly rounded) agree."""
    import torch

    from subsetix_amd.interval_stencil import build_stencil_plan
    from subsetix_amd.levels import gradient_tag

    npdt = np.float32 if dt == "f32" else np.float64
    tdt = torch.float32 if dt == "f32" else torch.float64
    for t in (0.02, 0.3, 1.7, 1e-30, 0.0, -1.0, np.inf, np.nan):
        gs = _ulp_neighbours(t, npdt) if np.isfinite(t) and t > 0 else [npdt(0), npdt(1e-20), npdt(0.5), npdt(3e38)]
        gs = [g for g in gs if np.isfinite(g)] + [npdt(np.inf)]
        nrow = len(gs)
        # row 2k holds [0, 2) with values (0, g_k); odd rows are empty
        off = [0]
        for k in range(2 * nrow):
            off.append(off[-1] + (1 if k % 2 == 0 else 0))
        s = ox.make(off, [0] * nrow, [2] * nrow)
        vals = np.zeros(2 * nrow, npdt)
        vals[1::2] = gs
        f = _field(sx, s, vals, tdt)
        exp = ox.grad_tag(s, vals, float(t))
        plan = build_stencil_plan(f.interval_set, f.interval_cell_offsets, 4)
        for p in (plan, None):
            got = to_host(gradient_tag(f, float(t), plan=p))
            print("threshold boundary", dt, t, "seg" if p is not None else "rows", flush=True)
            for name in ("rows", "off", "beg", "end"):
                np.testing.assert_array_equal(getattr(got, name), getattr(exp, name), err_msg=f"{t} {name}")
        if np.isfinite(t) and t > 0:
            # the boundary is crossed: some g tag, some do not
            tagged = np.diff(exp.off)[::2]
            assert 0 < tagged.sum() < nrow, t
