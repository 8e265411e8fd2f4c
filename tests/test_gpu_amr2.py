"""GPU parity: regrid (threshold / gradient tagging, grading), geometry,
gather / scatter, synchronisation and the two-level AMR step loop."""

import numpy as np
import pytest

from oracle import oracle as ox
from tests.golden_util import assert_set_equal, canonical_rows, runs_of, spans_set
from tests.gpu_util import to_device, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sx():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    import subsetix_amd

    return subsetix_amd


def _dense_field(sx, data):
    import torch

    h, w = data.shape
    s = sx.full_interval_set(w, h)
    f = sx.create_interval_field(s, 0.0, dtype=torch.float32)
    f.values.copy_(torch.from_numpy(np.ascontiguousarray(data, dtype=np.float32).reshape(-1)))
    return f


def test_regrid_golden(sx, golden):
    from subsetix_amd.amr2 import regrid
    from tests.test_oracle_golden import _dense_from_case

    for case in golden["regrid"]:
        print("regrid", case["id"], flush=True)
        if case["op"] == "grading":
            s = to_device(spans_set(case["height"], case["spans"]))
            got = regrid.enforce_two_level_grading_set(s, padding=case["padding"], mode=case["mode"],
                                                       width=case["width"], height=case["height"])
        else:
            f = _dense_field(sx, _dense_from_case(case))
            if case["op"] == "tag_value":
                got = regrid.gradient_tag_threshold_set(f, width=case["width"], height=case["height"],
                                                        threshold=case["threshold"])
            else:
                got = regrid.gradient_tag_set(f, width=case["width"], height=case["height"],
                                              frac_high=case["frac_high"], epsilon=case.get("epsilon", 1e-8))
        assert_set_equal(to_host(got), spans_set(case["height"], case["expect_spans"]))
    with pytest.raises(ValueError):
        regrid.enforce_two_level_grading_set(to_device(spans_set(4, {})), padding=1, mode="diag", width=4, height=4)


@pytest.mark.parametrize("wh", [(1, 5), (33, 7), (64, 3), (65, 4), (130, 9), (1000, 20)])
def test_threshold_and_gradient_vs_oracle(sx, wh):
    from subsetix_amd.amr2 import regrid

    w, h = wh
    rng = np.random.default_rng(w * 31 + h)
    d = rng.random((h, w)).astype(np.float32)
    d[d < 0.15] = np.nan
    f = _dense_field(sx, d)
    got = regrid.intervals_above_threshold_interval_field(f, width=w, height=h, threshold=0.5)
    assert_set_equal(to_host(got), ox.threshold(d, w, h, 0.5))
    d2 = np.nan_to_num(d, nan=0.0)
    f2 = _dense_field(sx, d2)
    got = regrid.gradient_tag_threshold_interval_field(f2, width=w, height=h, threshold=0.3)
    assert_set_equal(to_host(got), ox.gradient_tag(d2, w, h, 0.3))
    g = regrid.gradient_magnitude_interval_field(f2, width=w, height=h)
    np.testing.assert_array_equal(g.values.cpu().numpy(), ox.gradient(d2, w, h))
    full = np.ones((h, w), np.float32)
    got = regrid.intervals_above_threshold_interval_field(_dense_field(sx, full), width=w, height=h, threshold=0.5)
    assert canonical_rows(to_host(got)) == {y: [(0, w)] for y in range(h)}


def test_geometry_golden(sx, golden):
    from subsetix_amd.amr2 import ActionField, TwoLevelGeometry

    for case in golden["geometry"]:
        print("geometry", case["id"], flush=True)
        w, h, r = case["width"], case["height"], case["ratio"]
        actions = ActionField.full_grid(h, w, ratio=r)
        actions.set_from_interval_set(to_device(spans_set(h, case["refine_spans"])))
        geom = TwoLevelGeometry.from_action_field(actions)
        assert_set_equal(to_host(geom.coarse_only), spans_set(h, case["expect_coarse_only_spans"]))
        if case["expect_fine"] is not None:
            from tests.golden_util import oset

            assert_set_equal(to_host(geom.fine), oset(case["expect_fine"]))
        assert_set_equal(to_host(geom.refine), spans_set(h, case["refine_spans"]))
    # dilation modes (test_geometry.py:124-152)
    actions = ActionField.full_grid(5, 5, ratio=2)
    actions.set_from_interval_set(to_device(spans_set(5, {"2": [[2, 3]]})))
    geom = TwoLevelGeometry.from_action_field(actions)
    vn = geom.dilate_refine(halo=1, mode="von_neumann")
    assert_set_equal(to_host(vn.refine), spans_set(5, {"1": [[2, 3]], "2": [[1, 4]], "3": [[2, 3]]}))
    mo = geom.dilate_refine(halo=1, mode="moore")
    assert_set_equal(to_host(mo.refine), spans_set(5, {"1": [[1, 4]], "2": [[1, 4]], "3": [[1, 4]]}))


def test_action_field_refresh_from_values(sx):
    import torch

    from subsetix_amd.amr2 import Action, ActionField

    a = ActionField.full_grid(3, 6, ratio=2)
    v = a.values().view(3, 6)
    v[0, 1:4] = int(Action.REFINE)
    v[2, 0] = int(Action.REFINE)
    v[2, 5] = int(Action.REFINE)
    a._refine_set_cache = None
    got = to_host(a.refine_set())
    assert canonical_rows(got) == {0: [(1, 4)], 1: [], 2: [(0, 1), (5, 6)]}
    with pytest.raises(ValueError):
        a.set_from_interval_set(to_device(spans_set(3, {"1": [[4, 9]]})))
    assert a.values().dtype == torch.int8


def test_sync_and_gather_golden(sx, golden):
    import torch

    from subsetix_amd.amr2 import ActionField, gather_interval_subset, synchronize_interval_fields

    for case in golden["sync"]:
        print("sync", case["id"], flush=True)
        h, w = case["coarse_hw"]
        coarse = sx.create_interval_field(sx.full_interval_set(w, h), case["coarse_fill"], dtype=torch.float32)
        fine = sx.create_interval_field(sx.full_interval_set(2 * w, 2 * h), case.get("fine_fill", 0.0),
                                        dtype=torch.float32)
        if case.get("fine_arange"):
            fine.values.copy_(torch.arange(fine.values.numel(), dtype=torch.float32))
        fine0 = fine.values.clone()
        coarse0 = coarse.values.clone()
        actions = ActionField.full_grid(h, w, ratio=2)
        actions.set_from_interval_set(to_device(spans_set(h, case["refine_spans"])))
        copy = case["fill_fine_outside"]
        c2, f2 = synchronize_interval_fields(coarse, fine, actions, ratio=2, fill_fine_outside=case["fill_fine_outside"],
                                             copy=copy)
        if "expect_coarse_values" in case:
            assert c2 is coarse and f2 is fine
            np.testing.assert_allclose(c2.values.cpu().numpy(), np.array(case["expect_coarse_values"], np.float32))
            np.testing.assert_array_equal(f2.values.cpu().numpy(), fine0.cpu().numpy())
        else:
            exp_c = np.zeros((h, w), np.float32)
            for y, x in case["expect_coarse_cells_value_2"]:
                exp_c[y, x] = 2.0
            np.testing.assert_array_equal(c2.values.cpu().numpy().reshape(h, w), exp_c)
            y0, y1, x0, x1 = case["expect_fine_box_value_2"]
            exp_f = np.zeros((2 * h, 2 * w), np.float32)
            exp_f[y0:y1, x0:x1] = 2.0
            np.testing.assert_array_equal(f2.values.cpu().numpy().reshape(2 * h, 2 * w), exp_f)
            np.testing.assert_array_equal(coarse.values.cpu().numpy(), coarse0.cpu().numpy())
            np.testing.assert_array_equal(fine.values.cpu().numpy(), fine0.cpu().numpy())
    for case in golden["gather"]:
        h, w = case["super_hw"]
        f = sx.create_interval_field(sx.full_interval_set(w, h), 0.0, dtype=torch.float32)
        f.values.copy_(torch.arange(w * h, dtype=torch.float32))
        sub = gather_interval_subset(f, to_device(spans_set(h, case["subset_spans"])))
        np.testing.assert_array_equal(sub.values.cpu().numpy(), np.array(case["expect_values"], np.float32))
    with pytest.raises(TypeError):
        synchronize_interval_fields(coarse.values, fine, actions, ratio=2)
    with pytest.raises(ValueError):
        synchronize_interval_fields(coarse, fine, actions, ratio=3)


def _oracle_amr(u0, n, steps, regrid_every, vel, cfl, thr, grading, mode, tag, accumulate, dtype,
                return_refine=False):
    """simulation.py:232-354 restated with oracle primitives (full-domain
    levels, ratio 2)."""
    cv, fv, refine = ox.two_level_loop(u0, n, steps, regrid_every, velocity=vel, cfl=cfl, threshold=thr,
                                       padding=grading, mode=mode, tag=tag, accumulate=accumulate, dtype=dtype)
    return (cv, fv, refine) if return_refine else (cv, fv)


@pytest.mark.parametrize("dtype", ["f32", "f64"])
@pytest.mark.parametrize("accumulate", ["f32", "f64"])
@pytest.mark.parametrize("mode", ["von_neumann", "moore"])
@pytest.mark.parametrize("tag", ["gradient", "value"])
def test_amr_loop_vs_oracle(sx, dtype, accumulate, mode, tag):
    """The full step loop (sync, regrid, advance, sync) on a 32x32 Gaussian
    bump, bit-exact against the same loop on oracle primitives, for every
    regrid criterion (the reference default: gradient tag, von Neumann
    grading, FP32 sums) and option."""
    import torch

    from subsetix_amd.amr2 import AMR2Simulation, SimulationConfig

    n = 32
    xs = (np.arange(n) + 0.5) / n
    X, Y = np.meshgrid(xs, xs)
    u0 = np.exp(-((X - 0.3) ** 2 + (Y - 0.3) ** 2) / (2 * 0.1 ** 2))
    npdt = np.float32 if dtype == "f32" else np.float64
    tdt = torch.float32 if dtype == "f32" else torch.float64
    thr = 0.1 if tag == "value" else 0.02
    cfg = SimulationConfig(coarse_resolution=n, velocity=(1.0, 0.5), cfl=0.5, refine_threshold=thr, grading=2,
                           grading_mode=mode, tag=tag, accumulate=accumulate)
    sim = AMR2Simulation(cfg)
    field = sx.create_interval_field(sx.full_interval_set(n, n), 0.0, dtype=tdt)
    field.values.copy_(torch.from_numpy(u0.reshape(-1).astype(npdt)))
    sim.set_initial_field(field)
    sim.run(steps=12, regrid_every=5)
    cv, fv, refine = _oracle_amr(u0.reshape(-1), n, 12, 5, (1.0, 0.5), 0.5, thr, 2, mode, tag, accumulate, npdt,
                                 return_refine=True)
    assert 0 < refine.n
    np.testing.assert_array_equal(sim.state.coarse_field.values.cpu().numpy(), cv)
    np.testing.assert_array_equal(sim.state.fine_field.values.cpu().numpy(), fv)
    m = sim.total_mass()
    assert abs(m - ox.mass(cv, 1.0 / n / n)) <= 1e-12


def test_simulation_square_and_callbacks(sx):
    from subsetix_amd.amr2 import AMR2Simulation, SimulationConfig

    cfg = SimulationConfig(coarse_resolution=16, velocity=(0.1, 0.0), cfl=0.8, refine_threshold=0.05)
    sim = AMR2Simulation(cfg)
    sim.initialize_square()
    st = sim.state
    assert tuple(st.coarse.shape) == (16, 16) and tuple(st.fine.shape) == (32, 32)
    assert st.actions.refine_set().row_offsets.numel() == 17
    seen = []
    _, stats = sim.run(steps=2, listeners=[lambda s, t: seen.append(t.step)])
    assert stats.step == 2 and seen == [0, 1, 2]
    with pytest.raises(TypeError):
        sim.set_initial_field(st.coarse)
