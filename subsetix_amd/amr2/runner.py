"""Samurai-style two-level advection driver (drop-in for
``subsetix_amr2.runner``, runner.py:20-249): argument parsing, mesh set-up,
ghost update, snapshots and ``run_two_level_simulation``.

The time loop keeps the reference's order (adapt, clip the step to the end
time, update ghosts, advance both levels, swap) but advances the levels into
the second field's buffers with one stencil launch each instead of
allocating new fields every step; results are the same cell for cell.
"""

from __future__ import annotations

import argparse
import math
import os
from dataclasses import dataclass
from pathlib import Path
from typing import Callable, Sequence

from ..interval_field import IntervalField, create_interval_field
from ..interval_stencil import upwind_values
from .api import Box, MRAdaptor, ScalarField, TwoLevelMesh, make_scalar_field, synchronize_levels
from .export import save_two_level_vtk


@dataclass(frozen=True)
class OutputOptions:
    directory: Path
    prefix: str = "amr2"
    every: int = 10
    ghost_halo: int = 1

    def __post_init__(self) -> None:
        if self.every <= 0:
            raise ValueError(f"OutputOptions.every = {self.every}: snapshots need a positive period")


@dataclass(frozen=True)
class SimulationArgs:
    """Parameters of a two-level advection run (runner.py:37-55)."""

    min_corner: tuple[float, float] = (0.0, 0.0)
    max_corner: tuple[float, float] = (1.0, 1.0)
    velocity: tuple[float, float] = (1.0, 1.0)
    cfl: float = 0.5
    t0: float = 0.0
    tf: float = 0.1
    min_level: int = 4
    max_level: int = 5
    refine_threshold: float = 0.05
    grading: int = 1
    grading_mode: str = "von_neumann"
    output: OutputOptions | None = None


def _parser() -> argparse.ArgumentParser:
    p = argparse.ArgumentParser(description="two-level interval-set AMR advection (subsetix on MI355X)")
    pairs = (("--min-corner", (0.0, 0.0)), ("--max-corner", (1.0, 1.0)), ("--velocity", (1.0, 1.0)))
    for flag, default in pairs:
        p.add_argument(flag, type=float, nargs=2, default=default)
    p.add_argument("--cfl", type=float, default=0.5)
    p.add_argument("--Ti", dest="t0", type=float, default=0.0, help="initial time")
    p.add_argument("--Tf", dest="tf", type=float, default=0.1, help="final time")
    p.add_argument("--min-level", type=int, default=4)
    p.add_argument("--max-level", type=int, default=5)
    p.add_argument("--refine-threshold", type=float, default=0.05)
    p.add_argument("--grading", type=int, default=1)
    p.add_argument("--grading-mode", choices=["von_neumann", "moore"], default="von_neumann")
    p.add_argument("--vtk-dir", type=Path, default=None, help="write VTK snapshots here")
    p.add_argument("--vtk-prefix", default="amr2")
    p.add_argument("--vtk-every", type=int, default=10)
    p.add_argument("--ghost-halo", type=int, default=1)
    return p


def _check(args: SimulationArgs) -> SimulationArgs:
    if args.max_level != args.min_level + 1:
        raise ValueError(f"levels {args.min_level}..{args.max_level}: the driver runs exactly two adjacent levels")
    if args.tf < args.t0:
        raise ValueError(f"final time {args.tf} is before the initial time {args.t0}")
    if args.refine_threshold <= 0.0:
        raise ValueError(f"refine_threshold {args.refine_threshold} has to be positive")
    return args


def parse_simulation_args(argv: Sequence[str] | None = None) -> SimulationArgs:
    ns = _parser().parse_args(argv)
    out = None
    if ns.vtk_dir is not None:
        out = OutputOptions(directory=Path(ns.vtk_dir), prefix=str(ns.vtk_prefix), every=int(ns.vtk_every),
                            ghost_halo=int(ns.ghost_halo))
    pair = lambda v: (float(v[0]), float(v[1]))  # noqa: E731
    return _check(SimulationArgs(min_corner=pair(ns.min_corner), max_corner=pair(ns.max_corner),
                                 velocity=pair(ns.velocity), cfl=float(ns.cfl), t0=float(ns.t0), tf=float(ns.tf),
                                 min_level=int(ns.min_level), max_level=int(ns.max_level),
                                 refine_threshold=float(ns.refine_threshold), grading=int(ns.grading),
                                 grading_mode=ns.grading_mode, output=out))


def build_mesh(args: SimulationArgs, ratio: int = 2) -> TwoLevelMesh:
    """Coarse level of 2^min_level cells per side over the args' box."""
    return TwoLevelMesh(Box(args.min_corner, args.max_corner), 0, 1, ratio=ratio,
                        coarse_resolution=1 << int(args.min_level))


def update_ghost(field: ScalarField) -> None:
    """Synchronise the field's levels (restrict onto the refine set, parents
    onto the fine cells outside it)."""
    synchronize_levels(field)


def save_snapshot(output: OutputOptions, mesh: TwoLevelMesh, field: ScalarField, step: int, *,
                  time_value: float) -> None:
    g = mesh.geometry
    if g is None:
        raise RuntimeError("the mesh has no geometry yet")
    save_two_level_vtk(os.fspath(output.directory), output.prefix, step, coarse_field=field.coarse_field,
                       fine_field=field.fine_field, refine_set=g.refine, coarse_only_set=g.coarse_only,
                       fine_set=g.fine, dx_coarse=(mesh.box.max_corner[0] - mesh.box.min_corner[0]) / g.width,
                       dy_coarse=(mesh.box.max_corner[1] - mesh.box.min_corner[1]) / g.height, ratio=mesh.ratio,
                       time_value=time_value, ghost_halo=output.ghost_halo, width=g.width, height=g.height)


def _advance_into(src: IntervalField, dst: IntervalField, width: int, height: int, a: float, b: float, dt: float,
                  dx: float, dy: float) -> None:
    upwind_values(src, src.values, dst.values, width=width, height=height, a=a, b=b, dt=dt, dx=dx, dy=dy,
                  check_rows=False)


def run_two_level_simulation(args: SimulationArgs, init_fn: Callable[[ScalarField], None]) -> ScalarField:
    """Initialise with ``init_fn``, then advect to ``args.tf`` adapting the
    mesh every step (runner.py:172-249)."""
    mesh = build_mesh(args)
    field = make_scalar_field("u", mesh)
    init_fn(field)
    update_ghost(field)
    adapt = MRAdaptor(field, refine_threshold=args.refine_threshold, grading=args.grading, mode=args.grading_mode)
    adapt()
    g = mesh.geometry
    a, b = args.velocity
    dx0 = (args.max_corner[0] - args.min_corner[0]) / g.width
    dy0 = (args.max_corner[1] - args.min_corner[1]) / g.height
    dx1, dy1 = dx0 / mesh.ratio, dy0 / mesh.ratio
    rate = abs(a) / dx1 + abs(b) / dy1
    dt = args.cfl / rate if rate > 0 else max(args.tf - args.t0, 0.0)
    if dt <= 0.0:
        raise ValueError(f"time step {dt} from cfl={args.cfl}: the run would not advance")
    if args.output is not None:
        args.output.directory.mkdir(parents=True, exist_ok=True)
        save_snapshot(args.output, mesh, field, 0, time_value=args.t0)
    nxt = make_scalar_field("unp1", mesh)
    t, step = args.t0, 0
    while t < args.tf:
        adapt()
        h = min(dt, args.tf - t)
        update_ghost(field)
        g = mesh.geometry
        if field.coarse_field.values.dtype != nxt.dtype or nxt.coarse_field.interval_set.row_count != g.height:
            nxt = make_scalar_field("unp1", mesh, dtype=field.coarse_field.values.dtype)
        r = mesh.ratio
        _advance_into(field.coarse_field, nxt.coarse_field, g.width, g.height, a, b, h, dx0, dy0)
        _advance_into(field.fine_field, nxt.fine_field, g.width * r, g.height * r, a, b, h, dx1, dy1)
        field.swap(nxt)
        t = min(args.tf, t + h)
        step += 1
        if args.output is not None and (step % args.output.every == 0 or math.isclose(t, args.tf)):
            save_snapshot(args.output, mesh, field, step, time_value=t)
    return field


__all__ = ["OutputOptions", "SimulationArgs", "parse_simulation_args", "build_mesh", "update_ghost", "save_snapshot",
           "run_two_level_simulation"]
