"""SeaFloorConfig (src/alb_sim/config/sea_floor.py:4-16)."""
from dataclasses import dataclass

from ._meta import param


@dataclass(frozen=True)
class SeaFloorConfig:
    albedo: float = param(0.05, description="Proportion of light reflected")

    def __post_init__(self):
        if not 0 <= self.albedo <= 1:
            raise ValueError("Albedo must be between 0 and 1")
