"""SeaSurfaceConfig (src/alb_sim/config/sea_surface.py:4-19).  The GGX alpha
used by the model is roughness**2 (physics/reflection/microfacet_brdf.py:125)."""
from dataclasses import dataclass

from ._meta import param


@dataclass(frozen=True)
class SeaSurfaceConfig:
    albedo: float = param(0.1, description="Proportion of light reflected")
    roughness: float = param(0.05, description="Roughness of the surface")

    def __post_init__(self):
        if not 0 <= self.albedo <= 1:
            raise ValueError("Albedo must be between 0 and 1")
