"""Two-level AMR on interval sets (drop-in for ``subsetix_amr2``,
subsetix_amr2/__init__.py:9-62), on the gfx950 kernels of subsetix_amd."""

from .api import Box, MRAdaptor, ScalarField, TwoLevelMesh, make_scalar_field
from .export import save_two_level_vtk
from .fields import (
    Action,
    ActionField,
    clone_interval_field,
    gather_interval_subset,
    scatter_interval_subset,
    sync_levels,
    synchronize_interval_fields,
)
from .geometry import TwoLevelGeometry
from .mass import total_mass, total_mass_async
from .regrid import (
    enforce_two_level_grading_set,
    gradient_magnitude_interval_field,
    gradient_tag_set,
    gradient_tag_threshold_interval_field,
    gradient_tag_threshold_set,
    intervals_above_threshold_interval_field,
)
from .runner import (
    OutputOptions,
    SimulationArgs,
    build_mesh,
    parse_simulation_args,
    run_two_level_simulation,
    save_snapshot,
    update_ghost,
)
from .simulation import (
    AMR2Simulation,
    AMRState,
    SimulationConfig,
    SimulationStats,
    SquareSpec,
    TwoLevelVTKExporter,
    build_square_interval_field,
)

__all__ = [
    "TwoLevelGeometry",
    "gradient_tag_set",
    "gradient_tag_threshold_set",
    "enforce_two_level_grading_set",
    "synchronize_interval_fields",
    "save_two_level_vtk",
    "Box",
    "TwoLevelMesh",
    "make_scalar_field",
    "MRAdaptor",
    "OutputOptions",
    "SimulationArgs",
    "parse_simulation_args",
    "build_mesh",
    "run_two_level_simulation",
    "save_snapshot",
    "update_ghost",
    "AMR2Simulation",
    "AMRState",
    "SimulationConfig",
    "SimulationStats",
    "SquareSpec",
    "TwoLevelVTKExporter",
    "build_square_interval_field",
    # beyond the reference surface
    "Action",
    "ActionField",
    "ScalarField",
    "gradient_tag_threshold_interval_field",
    "gradient_magnitude_interval_field",
    "intervals_above_threshold_interval_field",
    "gather_interval_subset",
    "scatter_interval_subset",
    "clone_interval_field",
    "sync_levels",
    "total_mass",
    "total_mass_async",
]
