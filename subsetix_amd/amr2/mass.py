"""FP64 mass reduction: sum(u) * cell_area in double precision with a fixed
reduction order (4096-cell chunks, fixed tree), so the result does not
depend on the launch and is reproducible run to run."""

from __future__ import annotations

import torch

from .. import _lib
from .._lib import _F64, _I32, _I64, _P, check, context, ptr

_lib.register("stx_mass_parts", _I64, [_I64])
_lib.register("stx_mass", _I32, [_I32, _P, _I64, _F64, _P, _P, _P])


def total_mass_async(values: torch.Tensor, cell_area: float, out: torch.Tensor | None = None) -> torch.Tensor:
    """Device-side FP64 mass (a 1-element float64 tensor, no host sync)."""
    ctx = context(values.device)
    n = values.numel()
    parts = torch.empty(int(_lib.lib().stx_mass_parts(n)), dtype=torch.float64, device=values.device)
    if out is None:
        out = torch.empty(1, dtype=torch.float64, device=values.device)
    code = 4 if values.dtype == torch.float32 else 8
    if values.dtype not in (torch.float32, torch.float64):
        raise TypeError("mass needs float32 or float64 values")
    check(_lib.lib().stx_mass(code, ptr(values), n, float(cell_area), ptr(parts), ptr(out), ctx.stream), "mass")
    return out


def total_mass(values: torch.Tensor, cell_area: float) -> float:
    return float(total_mass_async(values, cell_area).item())
