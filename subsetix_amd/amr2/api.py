"""Samurai-style two-level wrappers (drop-in for ``subsetix_amr2.api``,
api.py:26-237): ``Box``, ``TwoLevelMesh``, ``ScalarField`` /
``make_scalar_field`` and the gradient ``MRAdaptor``.

The mesh owns the action field and geometry of the current refine set and,
next to them, the target of the ghost fill (the fine cells outside the
refine set), so ``update_ghost`` and the adaptor synchronise the two
full-domain levels with two fused transfer passes whose segment tables are
built once per regrid.
"""

from __future__ import annotations

from dataclasses import dataclass
from typing import Optional

import torch

from ..expressions import IntervalSet
from ..interval_field import IntervalField, create_interval_field, torch_dtype
from ..levels import prolong_into
from ..morphology import full_interval_set
from ..multilevel import prolong_set
from .fields import ActionField, sync_levels
from .geometry import TwoLevelGeometry
from .regrid import enforce_two_level_grading_set, gradient_tag_threshold_interval_field


@dataclass(frozen=True)
class Box:
    """Axis-aligned domain box."""

    min_corner: tuple[float, float]
    max_corner: tuple[float, float]


class TwoLevelMesh:
    """Levels 0 (coarse, N x N) and 1 (fine, ratio N x ratio N) over ``box``."""

    def __init__(self, box: Box, min_level: int, max_level: int, *, ratio: int = 2,
                 coarse_resolution: Optional[int] = None):
        if (int(min_level), int(max_level)) != (0, 1):
            raise ValueError(f"levels ({min_level}, {max_level}) requested; this mesh has exactly levels 0 and 1")
        self.box = box
        self.min_level, self.max_level = 0, 1
        self.ratio = int(ratio)
        self.geometry: Optional[TwoLevelGeometry] = None
        self.actions: Optional[ActionField] = None
        self.fine_outside: Optional[IntervalSet] = None
        self._resolution: Optional[int] = None
        if coarse_resolution is not None:
            if int(coarse_resolution) <= 0:
                raise ValueError(f"coarse_resolution {coarse_resolution} is not a positive cell count")
            self._resolution = int(coarse_resolution)
            self.initialise()

    def _install(self, width: int, height: int, refine: Optional[IntervalSet]) -> TwoLevelGeometry:
        actions = ActionField.full_grid(height, width, self.ratio)
        if refine is not None:
            actions.set_from_interval_set(refine)
        ws = self.geometry.workspace if self.geometry is not None else None
        self.actions = actions
        self.geometry = TwoLevelGeometry.from_action_field(actions, workspace=ws)
        self.fine_outside = prolong_set(self.geometry.coarse_only, self.ratio)
        return self.geometry

    def initialise(self, refine: IntervalSet | None = None) -> None:
        """Full coarse level, refined on ``refine`` when given."""
        if self._resolution is None:
            raise RuntimeError("initialise() needs the coarse_resolution given at construction")
        n = self._resolution
        if refine is not None:
            if not isinstance(refine, IntervalSet):
                raise TypeError(f"refine: expected an IntervalSet, got {type(refine).__name__}")
            if refine.row_count != n:
                raise ValueError(f"refine has {refine.row_count} rows, the mesh {n}")
        self._install(n, n, refine)

    def regrid(self, refine: IntervalSet) -> TwoLevelGeometry:
        """Replace the refine set (rows 0..height-1 of the coarse level)."""
        if not isinstance(refine, IntervalSet):
            raise TypeError(f"refine: expected an IntervalSet, got {type(refine).__name__}")
        rows = refine.row_count
        if self.geometry is None:
            if rows <= 0:
                raise ValueError("refine has no rows, so it cannot define the mesh")
            width = self._resolution or rows
            self._resolution = width
            return self._install(width, rows, refine)
        if rows != self.geometry.height:
            raise ValueError(f"refine has {rows} rows, the mesh {self.geometry.height}")
        return self._install(self.geometry.width, rows, refine)

    def cell_length(self, level: int) -> float:
        if self.geometry is None:
            raise RuntimeError("the mesh has no geometry yet")
        dx0 = (self.box.max_corner[0] - self.box.min_corner[0]) / float(self.geometry.width)
        return dx0 if level == self.min_level else dx0 / self.ratio


class ScalarField:
    """A scalar on both levels of a mesh: full-domain coarse and fine
    interval fields."""

    def __init__(self, name: str, mesh: TwoLevelMesh, *, dtype=torch.float32):
        if mesh.geometry is None:
            raise RuntimeError("create fields after the mesh is initialised")
        self.name = name
        self.mesh = mesh
        self.dtype = torch_dtype(dtype)
        self.coarse_field: IntervalField
        self.fine_field: IntervalField
        self.resize()

    def interval_field(self, level: int) -> IntervalField:
        return self.coarse_field if level == self.mesh.min_level else self.fine_field

    def set_interval_fields(self, *, coarse: IntervalField | None = None, fine: IntervalField | None = None) -> None:
        for what, f in (("coarse", coarse), ("fine", fine)):
            if f is not None and not isinstance(f, IntervalField):
                raise TypeError(f"{what}: expected an IntervalField, got {type(f).__name__}")
        if coarse is not None:
            self.coarse_field = coarse
        if fine is not None:
            self.fine_field = fine

    def resize(self) -> None:
        """Fresh zero fields on the mesh's current extent."""
        g = self.mesh.geometry
        if g is None:
            raise RuntimeError("the mesh has no geometry yet")
        w, h, r = g.width, g.height, self.mesh.ratio
        self.coarse_field = create_interval_field(full_interval_set(w, h), 0.0, dtype=self.dtype)
        self.fine_field = create_interval_field(full_interval_set(w * r, h * r), 0.0, dtype=self.dtype)

    def swap(self, other: "ScalarField") -> None:
        self.coarse_field, other.coarse_field = other.coarse_field, self.coarse_field
        self.fine_field, other.fine_field = other.fine_field, self.fine_field


def make_scalar_field(name: str, mesh: TwoLevelMesh, *, dtype=torch.float32) -> ScalarField:
    return ScalarField(name, mesh, dtype=dtype)


def synchronize_levels(field: ScalarField) -> None:
    """Coarse cells under the refine set take the mean of their children; fine
    cells outside it take their parent's value (in place)."""
    mesh = field.mesh
    if mesh.geometry is None or mesh.actions is None:
        raise RuntimeError("the mesh has no geometry yet")
    sync_levels(field.coarse_field, field.fine_field, mesh.actions.refine_set(), mesh.fine_outside,
                ratio=mesh.ratio)


class MRAdaptor:
    """Regrid from the coarse level's gradient: tag |grad u| >= threshold,
    grade, regrid the mesh, restart the fine level from its parents and
    synchronise (api.py:191-237)."""

    def __init__(self, field: ScalarField, *, refine_threshold: float, grading: int, mode: str = "von_neumann"):
        self.field = field
        self.mesh = field.mesh
        self.refine_threshold = float(refine_threshold)
        self.grading = int(grading)
        self.mode = mode

    def __call__(self) -> None:
        g = self.mesh.geometry
        if g is None:
            raise RuntimeError("the mesh has no geometry yet")
        f = self.field
        tags = gradient_tag_threshold_interval_field(f.coarse_field, width=g.width, height=g.height,
                                                     threshold=self.refine_threshold)
        graded = enforce_two_level_grading_set(tags, padding=self.grading, mode=self.mode, width=g.width,
                                               height=g.height)
        self.mesh.regrid(graded)
        # a new fine field (callers may hold the old one) on the same level set
        fine = create_interval_field(f.fine_field.interval_set, 0.0, dtype=f.coarse_field.values.dtype)
        prolong_into(fine.interval_set, fine, f.coarse_field, self.mesh.ratio)
        f.fine_field = fine
        synchronize_levels(f)


__all__ = ["Box", "TwoLevelMesh", "ScalarField", "make_scalar_field", "MRAdaptor", "synchronize_levels"]
