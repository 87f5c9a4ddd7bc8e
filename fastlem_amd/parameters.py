"""Per-site parameters: mirror of ``TopographicalParameters``
(src/core/parameters.rs:24-69), with the reference defaults
(base 0.0, erodibility 1.0, uplift 1.0, not an outlet, no max slope)."""
from __future__ import annotations

from dataclasses import dataclass, replace
from typing import Optional


@dataclass(frozen=True)
class TopographicalParameters:
    base_elevation: float = 0.0
    erodibility: float = 1.0
    uplift_rate: float = 1.0
    is_outlet: bool = False
    max_slope: Optional[float] = None  # radians, in [0, pi/2)

    # builder-style setters returning a new value (parameters.rs:44-69)
    def set_base_elevation(self, v: float) -> "TopographicalParameters":
        return replace(self, base_elevation=float(v))

    def set_erodibility(self, v: float) -> "TopographicalParameters":
        return replace(self, erodibility=float(v))

    def set_uplift_rate(self, v: float) -> "TopographicalParameters":
        return replace(self, uplift_rate=float(v))

    def set_is_outlet(self, v: bool) -> "TopographicalParameters":
        return replace(self, is_outlet=bool(v))

    def set_max_slope(self, v: Optional[float]) -> "TopographicalParameters":
        return replace(self, max_slope=None if v is None else float(v))
