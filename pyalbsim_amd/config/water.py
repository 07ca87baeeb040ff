"""Water column configuration (src/alb_sim/config/water.py:9-125)."""
import math
from dataclasses import dataclass
from typing import Optional, Tuple

from ..utils.parameter_profile import NumberOrScalar, normalize_number_or_scalar
from ._meta import param


@dataclass(frozen=True)
class FournierForandConfig:
    refractive_index_ratio: float = param(1.1, description="Refractive index ratio of particles relative to pure water (real part)")
    junge_slope: float = param(3.62, description="Slope of the hyperbolic (Junge) particle size distribution")


@dataclass(frozen=True)
class TurbidityLayerConfig:
    height: float = param(3, unit="m", description="Height of the turbidity layer")
    scattering_coefficient: NumberOrScalar = param(2.5, unit="1/m", description="Total scattering coefficient")
    absorption_coefficient: NumberOrScalar = param(0.169, unit="1/m", description="Total absorption coefficient")
    salinity: NumberOrScalar = param(37, unit="‰ (ppt)", description="Salt content of water (seawater is ~35, pure water is 0)")
    refractive_index: NumberOrScalar = param(1.33334, description="Refractive index of water")
    fournier_forand_parameters: FournierForandConfig = param(
        unit="FournierForandConfig", description="Fournier-Forand phase function parameters", factory=FournierForandConfig)
    sublayer_dz: Optional[float] = param(0.05, description="Height of the sublayers")
    num_sublayers: Optional[int] = param(None, description="Number of sublayers")

    def __post_init__(self):
        if self.height <= 0:
            raise ValueError("Layer height must be positive")
        for name in ("scattering_coefficient", "absorption_coefficient", "salinity", "refractive_index"):
            object.__setattr__(self, name, normalize_number_or_scalar(getattr(self, name)))
        if self.num_sublayers is None:
            if self.sublayer_dz is None:
                raise ValueError("Either sublayer_dz or num_sublayers must be provided")
            object.__setattr__(self, "num_sublayers", max(2, int(math.ceil(self.height / self.sublayer_dz))))


@dataclass(frozen=True)
class WaterConfig:
    layers: Tuple[TurbidityLayerConfig, ...] = param(
        unit="TurbidityLayerConfig", description="Tuple of turbidity layer configs",
        factory=lambda: (TurbidityLayerConfig(),))
