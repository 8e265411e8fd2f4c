"""GPU: the subsetix_amr2 facade (api.py / runner.py / export.py), point
location (locate_interval_cells, get_cell, set_cell), DLPack inputs, the
row-alignment helper and the benchmark suite.

* test_api.py:13-86 cases (mesh / field / adaptor, swap, storage
  replacement, dense inputs rejected);
* test_runner.py:23-48 known answer (first column 0.996, 1 elsewhere), and a
  32^2 Gaussian run of run_two_level_simulation against the runner loop
  restated on oracle primitives (runner.py:172-249);
* test_export.py cases plus the file read back: cell counts per (level,
  ghost) against the oracle's sets, values, coordinates;
* point queries against the oracle's cell model (absent rows, gaps,
  negative coordinates).
"""

import numpy as np
import pytest

from oracle import oracle as ox
from tests.golden_util import random_canonical
from tests.gpu_util import to_device, to_host

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def torch_mod():
    import torch

    if not torch.cuda.is_available():
        pytest.skip("no HIP device")
    return torch


# -------------------------------------------------------------------- api
def test_mesh_field_and_adaptor(torch_mod):
    torch = torch_mod
    from subsetix_amr2.api import Box, MRAdaptor, TwoLevelMesh, make_scalar_field

    mesh = TwoLevelMesh(Box((0.0, 0.0), (1.0, 1.0)), 0, 1, ratio=2, coarse_resolution=8)
    assert mesh.geometry is not None
    field = make_scalar_field("u", mesh)
    assert field.coarse_field.interval_set.row_count == 8 and field.coarse_field.values.numel() == 64
    assert field.fine_field.interval_set.row_count == 16 and field.fine_field.values.numel() == 256
    initial = torch.linspace(0.0, 1.0, 64, dtype=torch.float32, device="cuda")
    field.coarse_field.values[...] = initial
    MRAdaptor(field, refine_threshold=0.05, grading=1)()
    g = mesh.geometry
    assert g.refine.row_offsets.numel() == g.height + 1 and g.refine.begin.dtype == torch.int32
    assert abs(mesh.cell_length(0) - 1 / 8) < 1e-12 and abs(mesh.cell_length(1) - 1 / 16) < 1e-12
    # the adaptor against the oracle: gradient tag + von Neumann grading
    d = initial.cpu().numpy().reshape(8, 8)
    exp = ox.grading(ox.gradient_tag(d, 8, 8, 0.05), padding=1, mode="von_neumann", width=8, height=8)
    assert ox.cells(to_host(g.refine)) == ox.cells(exp) and ox.cells(exp)
    # fine level: parents outside the refine set, coarse = mean of children inside
    fine = field.fine_field.values.cpu().numpy().reshape(16, 16)
    np.testing.assert_array_equal(fine, np.kron(d, np.ones((2, 2), np.float32)))


def test_scalar_field_swap_and_storage(torch_mod):
    torch = torch_mod
    from subsetix_amr2.api import Box, TwoLevelMesh, make_scalar_field
    from subsetix_cupy import create_interval_field

    mesh = TwoLevelMesh(Box((0.0, 0.0), (1.0, 1.0)), 0, 1, coarse_resolution=4)
    u, v = make_scalar_field("u", mesh), make_scalar_field("v", mesh)
    u.coarse_field.values.fill_(1.0)
    v.coarse_field.values.fill_(2.0)
    u.swap(v)
    assert torch.all(u.coarse_field.values == 2.0) and torch.all(v.coarse_field.values == 1.0)
    m2 = TwoLevelMesh(Box((0.0, 0.0), (1.0, 1.0)), 0, 1, coarse_resolution=2)
    f = make_scalar_field("u", m2)
    new = create_interval_field(f.coarse_field.interval_set, fill_value=3.0, dtype=torch.float32)
    f.set_interval_fields(coarse=new)
    assert f.coarse_field is new and torch.all(f.coarse_field.values == 3.0)
    dense = torch.zeros((4, 4), dtype=torch.int8, device="cuda")
    with pytest.raises(TypeError):
        mesh.initialise(refine=dense)
    with pytest.raises(TypeError):
        mesh.regrid(dense)
    with pytest.raises(TypeError):
        f.set_interval_fields(coarse=torch.zeros_like(f.coarse_field.values))
    with pytest.raises(ValueError):
        TwoLevelMesh(Box((0.0, 0.0), (1.0, 1.0)), 1, 3, coarse_resolution=4)
    with pytest.raises(ValueError):
        mesh.regrid(to_device(ox.make([0, 0, 0], [], [])))  # 2 rows on a 4-row mesh


# ----------------------------------------------------------------- runner
def test_runner_known_answer(torch_mod, golden):
    from subsetix_amr2.runner import SimulationArgs, run_two_level_simulation

    ka = golden["runner_known_answer"]

    def init_fn(f):
        f.coarse_field.values.fill_(1.0)
        f.fine_field.values.fill_(1.0)

    field = run_two_level_simulation(SimulationArgs(min_corner=(0.0, 0.0), max_corner=(1.0, 1.0), velocity=(0.1, 0.0),
                                                    cfl=0.5, t0=0.0, tf=0.01, min_level=2, max_level=3,
                                                    refine_threshold=0.05, grading=1), init_fn)
    g = field.mesh.geometry
    coarse = field.coarse_field.values.cpu().numpy().reshape(g.height, g.width)
    exp = np.full(coarse.shape, ka["expect_elsewhere"], np.float32)
    exp[:, 0] = ka["expect_first_column"]
    np.testing.assert_allclose(coarse, exp, atol=ka["atol"])


def _oracle_runner(cv, fv, n, vel, cfl, tf, thr, grading, mode):
    """runner.py:172-249 on oracle primitives (ratio 2, unit box)."""
    coarse, fine = ox.full(n, n), ox.full(2 * n, 2 * n)
    empty = ox.make([0] * (n + 1), [], [])

    def adapt(cv, fv):
        tags = ox.gradient_tag(cv.reshape(n, n), n, n, thr)
        refine = ox.grading(tags, padding=grading, mode=mode, width=n, height=n)
        _, fv = ox.prolong_field(coarse, cv, 2)
        cv, fv = ox.synchronize(coarse, cv, fine, fv, refine, ratio=2)
        return cv, fv, refine

    cv, fv = ox.synchronize(coarse, cv, fine, fv, empty, ratio=2)
    cv, fv, refine = adapt(cv, fv)
    dx0 = 1.0 / n
    dx1 = dx0 / 2
    dt = cfl / (abs(vel[0]) / dx1 + abs(vel[1]) / dx1)
    t, steps = 0.0, 0
    while t < tf:
        cv, fv, refine = adapt(cv, fv)
        h = min(dt, tf - t)
        cv, fv = ox.synchronize(coarse, cv, fine, fv, refine, ratio=2)
        cv = ox.upwind(coarse, cv, width=n, height=n, a=vel[0], b=vel[1], dt=h, dx=dx0, dy=dx0)
        fv = ox.upwind(fine, fv, width=2 * n, height=2 * n, a=vel[0], b=vel[1], dt=h, dx=dx1, dy=dx1)
        t = min(tf, t + h)
        steps += 1
    return cv, fv, refine, steps


def test_runner_vs_oracle(torch_mod):
    torch = torch_mod
    from subsetix_amd import workloads
    from subsetix_amr2.runner import SimulationArgs, run_two_level_simulation

    n = 32
    u0 = workloads.gaussian_bump(n, center=(0.35, 0.4), sigma=0.08).astype(np.float32).reshape(-1)
    f0 = np.random.default_rng(3).random(4 * n * n).astype(np.float32)

    def init_fn(f):
        f.coarse_field.values.copy_(torch.from_numpy(u0))
        f.fine_field.values.copy_(torch.from_numpy(f0))

    args = SimulationArgs(velocity=(1.0, 0.5), cfl=0.5, tf=0.05, min_level=5, max_level=6, refine_threshold=0.02,
                          grading=1, grading_mode="von_neumann")
    field = run_two_level_simulation(args, init_fn)
    cv, fv, refine, steps = _oracle_runner(u0, f0, n, (1.0, 0.5), 0.5, 0.05, 0.02, 1, "von_neumann")
    assert steps >= 9 and refine.n > 0
    np.testing.assert_array_equal(field.coarse_field.values.cpu().numpy(), cv)
    np.testing.assert_array_equal(field.fine_field.values.cpu().numpy(), fv)
    assert ox.cells(to_host(field.mesh.geometry.refine)) == ox.cells(refine)


# ----------------------------------------------------------------- export
def test_save_two_level_vtk(torch_mod, tmp_path):
    torch = torch_mod
    import subsetix_amd as sx
    from subsetix_amd.export_vtk import read_vtu
    from subsetix_amr2 import Action, ActionField, TwoLevelGeometry, save_two_level_vtk

    refine_h = ox.make([0, 0, 1, 2, 2], [1, 1], [3, 3])
    actions = ActionField.full_grid(4, 4, ratio=2, default=Action.KEEP)
    actions.set_from_interval_set(to_device(refine_h))
    geom = TwoLevelGeometry.from_action_field(actions)
    cf = sx.create_interval_field(sx.full_interval_set(4, 4), 1.0, dtype=torch.float32)
    ff = sx.create_interval_field(sx.full_interval_set(8, 8), 2.0, dtype=torch.float32)
    ff.values.copy_(torch.arange(64, dtype=torch.float32))
    files = save_two_level_vtk(str(tmp_path), "demo", 5, coarse_field=cf, fine_field=ff, refine_set=geom.refine,
                               coarse_only_set=geom.coarse_only, fine_set=geom.fine, dx_coarse=0.25, dy_coarse=0.25,
                               ratio=2, ghost_halo=1, width=4, height=4)
    assert files == {"mesh_vtu": "demo_mesh_step0005.vtu"}
    arrs = read_vtu(str(tmp_path / files["mesh_vtu"]))
    co = ox.difference(ox.full(4, 4), refine_h)
    cghost = ox.intersection(ox.ghost_zones(refine_h, 1, 1, width=4, height=4), co)
    cact = ox.difference(co, cghost)
    fine_h = ox.prolong_set(refine_h, 2)
    fghost = ox.ghost_zones(fine_h, 2, 2, width=8, height=8)
    groups = [(0, 0, cact), (0, 1, cghost), (1, 0, fine_h), (1, 1, fghost)]
    k = 0
    for level, ghost, s in groups:
        cells = sorted(ox.cells(s))  # storage order = (row, x) order
        n = len(cells)
        assert np.all(arrs["level"][k:k + n] == level) and np.all(arrs["ghost_mask"][k:k + n] == ghost)
        for j, (y, x) in enumerate(cells):
            exp_u = 1.0 if level == 0 else float(y * 8 + x)
            assert arrs["u"][k + j] == exp_u
            dx = 0.25 if level == 0 else 0.125
            np.testing.assert_allclose(arrs["points"][4 * (k + j)], [x * dx, y * dx, 0.0], rtol=0, atol=1e-7)
        k += n
    assert k == arrs["level"].size
    with pytest.raises(TypeError):
        save_two_level_vtk(str(tmp_path), "demo", 0, coarse_field=cf.values, fine_field=ff, refine_set=geom.refine,
                           coarse_only_set=geom.coarse_only, fine_set=geom.fine, dx_coarse=1.0, dy_coarse=1.0)


def test_vtk_exporter_callback(torch_mod, tmp_path):
    from subsetix_amr2 import AMR2Simulation, SimulationConfig, TwoLevelVTKExporter

    sim = AMR2Simulation(SimulationConfig(coarse_resolution=16, velocity=(0.1, 0.0), cfl=0.8, refine_threshold=0.05))
    sim.initialize_square()
    sim.run(steps=2, exporters=[TwoLevelVTKExporter(str(tmp_path), every=1, ghost_halo=1)])
    names = sorted(p.name for p in tmp_path.iterdir())
    assert names == [f"amr2_mesh_step{k:04d}.vtu" for k in range(3)]


# ---------------------------------------------------- point location (K9)
class _Producer:
    """A DLPack producer that is not a torch tensor (as a CuPy array is)."""

    def __init__(self, t):
        self._t = t
        self.dtype = np.dtype(str(t.dtype).replace("torch.", ""))
        self.shape = tuple(t.shape)

    def __dlpack__(self, stream=None):
        return self._t.__dlpack__()

    def __dlpack_device__(self):
        return self._t.__dlpack_device__()


def test_locate_interval_cells_vs_oracle(torch_mod):
    torch = torch_mod
    import subsetix_amd as sx

    rng = np.random.default_rng(4)
    s = random_canonical(rng, 40, 90, mean_n=4, rows=np.arange(-5, 75, 2), touching=True)
    f = sx.create_interval_field(to_device(s), 0.0, dtype=torch.float32)
    f.values.copy_(torch.arange(f.values.numel(), dtype=torch.float32))
    qr = rng.integers(-8, 80, size=(30, 17)).astype(np.int32)
    qc = rng.integers(-3, 95, size=(30, 17)).astype(np.int32)
    got = sx.locate_interval_cells(f, qr, qc).cpu().numpy()
    ir = ox.interval_rows(s)
    exp = np.full(qr.shape, -1, np.int32)
    for i in range(s.n):
        hit = (qr == ir[i]) & (qc >= s.beg[i]) & (qc < s.end[i])
        exp[hit & (exp < 0)] = i
    assert got.shape == qr.shape and got.dtype == np.int32
    np.testing.assert_array_equal(got, exp)
    assert (exp >= 0).sum() > 20 and (exp < 0).sum() > 20
    # DLPack producers and a caller buffer
    out = torch.empty(qr.size, dtype=torch.int32, device="cuda")
    res = sx.locate_interval_cells(f, _Producer(torch.from_numpy(qr).cuda()), _Producer(torch.from_numpy(qc).cuda()),
                                   out=_Producer(out))
    np.testing.assert_array_equal(res.cpu().numpy(), exp)
    np.testing.assert_array_equal(out.cpu().numpy().reshape(qr.shape), exp)
    with pytest.raises(ValueError):
        sx.locate_interval_cells(f, qr, qc[:, :3])
    assert sx.locate_interval_cells(f, np.zeros(0, np.int32), np.zeros(0, np.int32)).numel() == 0
    # get_cell / set_cell: value of the cell, None for gaps, IndexError for absent rows
    co = ox.cell_offsets(s)
    for (y, x), k in list(np.ndenumerate(exp))[:60]:
        row, col = int(qr[y, x]), int(qc[y, x])
        if row not in set(s.rows.tolist()):
            with pytest.raises(IndexError):
                sx.get_cell(f, row, col)
            continue
        v = sx.get_cell(f, row, col)
        if k < 0:
            assert v is None and sx.set_cell(f, row, col, -1.0) is False
        else:
            idx = int(co[k] + col - s.beg[k])
            assert float(v) == float(idx)
            assert sx.set_cell(f, row, col, -7.5) is True
            assert float(f.values[idx]) == -7.5


def test_dlpack_inputs_build_sets_and_fields(torch_mod):
    """DLPack producers (not torch tensors) as CSR arrays, as stencil output
    buffer and as query arrays, as CuPy arrays are handed over."""
    torch = torch_mod
    import subsetix_amd as sx

    s = ox.make([0, 2, 3, 5], [0, 5, 1, 0, 4], [3, 9, 8, 2, 9])
    t = lambda a: _Producer(torch.from_numpy(np.asarray(a, np.int32)).cuda())  # noqa: E731
    d = sx.IntervalSet(begin=t(s.beg), end=t(s.end), row_offsets=t(s.off), rows=t(s.rows))
    assert d.interval_count == 5 and d.row_count == 3
    e = sx.build_interval_set(t(s.off), t(s.beg), t(s.end))
    u = sx.evaluate(sx.make_union(sx.make_input(d), sx.make_input(e)))
    exp = ox.union(s, s)
    np.testing.assert_array_equal(u.begin.cpu().numpy(), exp.beg)
    f = sx.create_interval_field(d, 0.0, dtype=torch.float32)
    vals = np.linspace(0, 1, f.values.numel()).astype(np.float32)
    f.values.copy_(torch.from_numpy(vals))
    out = torch.empty_like(f.values)
    res = sx.step_upwind_interval(f, width=10, height=3, a=0.5, b=-0.25, dt=0.1, dx=0.1, dy=0.1, out=_Producer(out))
    ref = ox.upwind(s, vals, width=10, height=3, a=0.5, b=-0.25, dt=0.1, dx=0.1, dy=0.1)
    np.testing.assert_array_equal(out.cpu().numpy(), ref)
    np.testing.assert_array_equal(res.cpu().numpy(), ref)
    back = torch.from_dlpack(res)  # results leave through DLPack without a copy
    assert back.data_ptr() == out.data_ptr()


def test_align_interval_sets_vs_oracle(torch_mod):
    from subsetix_amd.expressions import _align_interval_sets

    rng = np.random.default_rng(6)
    a = random_canonical(rng, 20, 60, mean_n=3, rows=np.sort(rng.choice(50, 20, replace=False)))
    b = random_canonical(rng, 15, 60, mean_n=3, rows=np.sort(rng.choice(50, 15, replace=False)))
    la, lb, rows = _align_interval_sets(to_device(a), to_device(b))
    u = np.union1d(a.rows, b.rows)
    np.testing.assert_array_equal(rows.cpu().numpy(), u)
    for src, got in ((a, la), (b, lb)):
        h = to_host(got)
        np.testing.assert_array_equal(h.rows, u)
        np.testing.assert_array_equal(h.beg, src.beg)
        assert ox.cells(h) == ox.cells(src)
        np.testing.assert_array_equal(np.diff(h.off)[np.isin(u, src.rows)], np.diff(src.off))
        assert np.all(np.diff(h.off)[~np.isin(u, src.rows)] == 0)


def test_benchmark_suite_small_cases(torch_mod):
    """Every small case of the suite builds and runs (correctness of the
    operations themselves is tested elsewhere); ns/int is reported."""
    from subsetix_amd.benchmarks.__main__ import main

    recs = main(["--pattern", "_small$", "--repeat", "2", "--warmup", "1"])
    assert len(recs) == 24
    for r in recs:
        assert r["avg_ms"] > 0 and r["input_intervals"] > 0 and r["ns_per_interval"] > 0, r["name"]
