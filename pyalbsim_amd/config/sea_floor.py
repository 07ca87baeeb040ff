"""SeaFloorConfig (src/alb_sim/config/sea_floor.py:4-16).

slope / azimuth extend the reference's flat floor: the floor plane is tilted
by slope degrees about the beam's floor point, its normal leaning toward
azimuth degrees in the x-z plane (0 = +x, 90 = +z).  The defaults keep the
reference's flat floor.
"""
from dataclasses import dataclass

from ._meta import param


@dataclass(frozen=True)
class SeaFloorConfig:
    albedo: float = param(0.05, description="Proportion of light reflected")
    slope: float = param(0.0, unit="deg", description="Tilt of the floor plane (extension; 0 = flat)")
    azimuth: float = param(0.0, unit="deg", description="Direction of the tilt in the x-z plane (extension)")

    def __post_init__(self):
        if not 0 <= self.albedo <= 1:
            raise ValueError("Albedo must be between 0 and 1")
        if not 0 <= self.slope < 90:
            raise ValueError("Floor slope must be in [0, 90) degrees")
