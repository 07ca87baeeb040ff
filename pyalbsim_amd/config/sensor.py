"""SensorConfig (src/alb_sim/config/sensor.py:4-33)."""
from dataclasses import dataclass

from ._meta import param


@dataclass(frozen=True)
class SensorConfig:
    aperture_radius: float = param(0.1, unit="m", description="Receiver telescope radius")
    field_of_view: float = param(5e-3, unit="rad", description="Half-angle of field of view")
    sample_rate: int = param(int(2e9), unit="Hz", description="Temporal resolution of simulated waveform")

    def __post_init__(self):
        if self.aperture_radius < 0:
            raise ValueError("Aperture radius must be positive")
