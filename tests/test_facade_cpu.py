"""CPU: runner argument parsing (test_runner.py:8-16), the benchmark-suite
generators against per-row loop restatements of the reference generators
(benchmarks/utils.py:41-232), the VTU writer's file layout, and the
drop-in import names."""

import numpy as np
import pytest

from subsetix_amd.benchmarks import geometry as geo


def test_parse_simulation_args_defaults_and_checks():
    from subsetix_amd.amr2.runner import parse_simulation_args

    a = parse_simulation_args([])
    assert a.min_corner == (0.0, 0.0) and a.max_corner == (1.0, 1.0) and a.velocity == (1.0, 1.0)
    assert (a.min_level, a.max_level) == (4, 5) and a.output is None
    assert (a.cfl, a.t0, a.tf, a.grading, a.grading_mode) == (0.5, 0.0, 0.1, 1, "von_neumann")
    b = parse_simulation_args(["--velocity", "0.1", "0", "--Tf", "0.01", "--min-level", "2", "--max-level", "3",
                               "--vtk-dir", "/tmp/x", "--vtk-every", "3", "--grading-mode", "moore"])
    assert b.velocity == (0.1, 0.0) and b.tf == 0.01 and b.min_level == 2 and b.grading_mode == "moore"
    assert b.output is not None and b.output.every == 3 and str(b.output.directory) == "/tmp/x"
    for bad in (["--min-level", "4", "--max-level", "6"], ["--Ti", "1", "--Tf", "0.5"], ["--refine-threshold", "0"]):
        with pytest.raises(ValueError):
            parse_simulation_args(bad)
    from subsetix_amd.amr2.runner import OutputOptions

    with pytest.raises(ValueError):
        OutputOptions(directory="/tmp", every=0)


def _loop_square_grid(spec, shift_x=0, shift_y=0):
    """utils.py:103-182 restated: mask of squares, rolled, runs per row."""
    rows, width = spec.rows, spec.width
    tx, ty = max(1, min(4, spec.tiles_x)), max(1, spec.tiles_y)
    fill = float(np.clip(spec.fill_ratio, 0.05, 1.0))
    mask = np.zeros((rows, width), bool)
    ch, cw = rows / ty, width / tx
    sh, sw = max(1, int(round(ch * fill))), max(1, int(round(cw * fill)))
    for gy in range(ty):
        yc = (gy + 0.5) * ch
        y0, y1 = max(0, min(rows, int(round(yc - sh / 2)))), max(0, min(rows, int(round(yc + sh / 2))))
        for gx in range(tx):
            xc = (gx + 0.5) * cw
            x0, x1 = max(0, min(width, int(round(xc - sw / 2)))), max(0, min(width, int(round(xc + sw / 2))))
            if y1 > y0 and x1 > x0:
                mask[y0:y1, x0:x1] = True
    mask = np.roll(np.roll(mask, shift_y, axis=0), shift_x, axis=1)
    off, b, e = [0], [], []
    for row in mask:
        x = 0
        while x < width:
            if row[x]:
                s = x
                while x < width and row[x]:
                    x += 1
                b.append(s)
                e.append(x)
            else:
                x += 1
        off.append(len(b))
    return np.array(off), np.array(b), np.array(e)


@pytest.mark.parametrize("spec,shift", [(geo.GeometrySpec(64, 64, 3, 3, 0.8), (0, 0)),
                                        (geo.GeometrySpec(64, 96, 3, 3, 0.8), (16, 0)),
                                        (geo.GeometrySpec(90, 70, 5, 2, 0.6), (12, 15)),
                                        (geo.GeometrySpec(32, 32, 1, 1, 0.65), (0, 0))])
def test_square_grid_matches_loop_restatement(spec, shift):
    got = geo.square_grid_csr(spec, *shift)
    exp = _loop_square_grid(spec, *shift)
    for g, e in zip(got, exp):
        np.testing.assert_array_equal(g, e)
    assert got[1].size > 0


def test_random_set_matches_loop_restatement():
    """utils.py:41-100: same generator calls per row, disjoint sorted
    intervals inside [0, width)."""
    for rows, width, ipr, seed in ((40, 64, (1, 6), 123), (25, 9, (1, 9), 987), (10, 1, 3, 5)):
        off, b, e = geo.random_csr(rows=rows, width=width, intervals_per_row=ipr, seed=seed)
        rng = np.random.default_rng(seed)
        lo, hi = (ipr, ipr) if isinstance(ipr, int) else ipr
        k0 = 0
        for r in range(rows):
            count = int(rng.integers(max(1, lo), max(max(1, lo), hi) + 1))
            if width <= 1:
                count = 0
            count = min(count, max(0, width // 2))
            exp_b, exp_e = [], []
            if count:
                pts = np.sort(rng.choice(np.arange(0, width + 1, dtype=np.int32), size=2 * count, replace=False))
                last = 0
                for s, t in zip(pts[0::2], np.clip(np.maximum(pts[1::2], pts[0::2] + 1), 0, width)):
                    s = int(max(s, last))
                    if s >= width:
                        break
                    t = int(max(t, s + 1))
                    t = min(t, width)
                    if t <= s:
                        continue
                    exp_b.append(s)
                    exp_e.append(t)
                    last = t
            assert list(b[off[r]:off[r + 1]]) == exp_b and list(e[off[r]:off[r + 1]]) == exp_e
            k0 += len(exp_b)
        assert off[-1] == k0 and np.all(e > b) and np.all(e <= width)


def test_staircase_keeps_rows_and_lengths():
    spec = geo.GeometrySpec(64, 64, 1, 1, 0.65)
    off, b, e = geo.square_grid_csr(spec)
    so, sb, se = geo.staircase_csr(spec.width, off, b, e, steps=8)
    np.testing.assert_array_equal(so, off)
    np.testing.assert_array_equal(se - sb, e - b)
    assert np.all(sb >= 0) and np.all(se <= spec.width)
    starts = sb[off[:-1][np.diff(off) > 0]]  # each active row's first begin: its step's offset
    assert starts[0] == 0 and starts.max() > 0 and np.all(np.diff(starts) >= 0)
    # steps of height active_rows // 8 (a last partial step is clamped at the width)
    assert len(np.unique(starts)) >= 8 and starts.max() <= spec.width - (e.max() - b.min())


def test_vtu_writer_layout_and_roundtrip(tmp_path):
    from subsetix_amd.export_vtk import read_vtu, write_unstructured_quads_vtu

    mask = np.zeros((3, 4), bool)
    mask[1, 1:3] = True
    mask[2, 0] = True
    u = np.arange(12, dtype=np.float32).reshape(3, 4)
    ghost = np.zeros((3, 4), np.int32)
    ghost[2, 0] = 1
    path = tmp_path / "m.vtu"
    write_unstructured_quads_vtu(str(path), [(mask, u, 1, 0.5, 0.25, 0.0, 0.0, ghost)])
    text = path.read_bytes()
    for needle in (b'header_type="UInt64"', b'Name="ghost_mask"', b'Name="level"', b'NumberOfCells="3"',
                   b'NumberOfPoints="12"', b'<AppendedData encoding="raw">'):
        assert needle in text
    arrs = read_vtu(str(path))
    np.testing.assert_array_equal(arrs["level"], [1, 1, 1])
    np.testing.assert_array_equal(arrs["ghost_mask"], [0, 0, 1])
    np.testing.assert_array_equal(arrs["u"], [5, 6, 8])
    np.testing.assert_array_equal(arrs["types"], [9, 9, 9])
    np.testing.assert_array_equal(arrs["offsets"], [4, 8, 12])
    np.testing.assert_array_equal(arrs["connectivity"], np.arange(12))
    # first quad: cell (row 1, col 1): x in [0.5, 1.0], y in [0.25, 0.5], counter-clockwise
    np.testing.assert_allclose(arrs["points"][:4], [[0.5, 0.25, 0], [1.0, 0.25, 0], [1.0, 0.5, 0], [0.5, 0.5, 0]])
    empty = tmp_path / "e.vtu"
    write_unstructured_quads_vtu(str(empty), [])
    assert read_vtu(str(empty))["u"].size == 0


def test_drop_in_import_names():
    import subsetix_amr2
    import subsetix_cupy
    from subsetix_amr2.api import Box, MRAdaptor, TwoLevelMesh, make_scalar_field  # noqa: F401
    from subsetix_amr2.runner import SimulationArgs, run_two_level_simulation  # noqa: F401
    from subsetix_cupy.expressions import IntervalSet, evaluate  # noqa: F401
    from subsetix_cupy.interval_field import locate_interval_cells  # noqa: F401

    ref_cupy = ["Expr", "IntervalSet", "CuPyWorkspace", "build_interval_set", "evaluate", "make_difference",
                "make_complement", "make_input", "make_intersection", "make_symmetric_difference", "make_union",
                "IntervalField", "create_interval_field", "get_cell", "set_cell", "locate_interval_cells",
                "MultiLevel2D", "MultiLevelField2D", "prolong_set", "restrict_set", "prolong_field", "restrict_field",
                "prolong_level_sets", "restrict_level_sets", "prolong_level_field", "restrict_level_field",
                "covered_by_fine", "coarse_only", "translate_interval_set", "interior_for_direction",
                "boundary_for_direction", "boundary_layer", "dilate_interval_set", "erode_interval_set",
                "ghost_zones", "full_interval_set", "clip_interval_set", "step_upwind_interval",
                "step_upwind_interval_field"]
    ref_amr2 = ["TwoLevelGeometry", "gradient_tag_set", "gradient_tag_threshold_set", "enforce_two_level_grading_set",
                "synchronize_interval_fields", "save_two_level_vtk", "Box", "TwoLevelMesh", "make_scalar_field",
                "MRAdaptor", "OutputOptions", "SimulationArgs", "parse_simulation_args", "build_mesh",
                "run_two_level_simulation", "save_snapshot", "update_ghost", "AMR2Simulation", "AMRState",
                "SimulationConfig", "SimulationStats", "SquareSpec", "TwoLevelVTKExporter",
                "build_square_interval_field"]
    assert not [n for n in ref_cupy if not hasattr(subsetix_cupy, n)]
    assert not [n for n in ref_amr2 if not hasattr(subsetix_amr2, n)]
