"""Two-level AMR on interval sets (drop-in for ``subsetix_amr2``,
subsetix_amr2/__init__.py:9-62), on the gfx950 kernels of subsetix_amd."""

from .fields import (
    Action,
    ActionField,
    clone_interval_field,
    gather_interval_subset,
    scatter_interval_subset,
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
from .simulation import (
    AMR2Simulation,
    AMRState,
    SimulationConfig,
    SimulationStats,
    SquareSpec,
    build_square_interval_field,
)

__all__ = [
    "Action",
    "ActionField",
    "TwoLevelGeometry",
    "gradient_tag_set",
    "gradient_tag_threshold_set",
    "gradient_tag_threshold_interval_field",
    "gradient_magnitude_interval_field",
    "intervals_above_threshold_interval_field",
    "enforce_two_level_grading_set",
    "synchronize_interval_fields",
    "gather_interval_subset",
    "scatter_interval_subset",
    "clone_interval_field",
    "total_mass",
    "total_mass_async",
    "AMR2Simulation",
    "AMRState",
    "SimulationConfig",
    "SimulationStats",
    "SquareSpec",
    "build_square_interval_field",
]
