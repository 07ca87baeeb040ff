"""LaserConfig (src/alb_sim/config/laser.py:4-47).  ``nadir_angle`` and
``azimuth_angle`` are used in degrees by the model although the metadata says
rad (physics/models/laser.py:22-33); kept as is for compatibility."""
from dataclasses import dataclass

from ._meta import param


@dataclass(frozen=True)
class LaserConfig:
    divergence_angle: float = param(1e-3, unit="rad", description="Laser beam divergence half-angle")
    wavelength: float = param(0.532, unit="µm", description="Laser wavelength in vacuum")
    nadir_angle: float = param(15.0, unit="rad", description="Nadir angle from -y axis")
    azimuth_angle: float = param(0.0, unit="rad", description="Azimuth angle around y axis (0 = +x)")
    pulse_fwhm: float = param(8.3e-9, unit="s", description="Full width at half maximum of laser pulse")

    def __post_init__(self):
        if self.nadir_angle >= 90:
            raise ValueError("Nadir angle must be below 90 degrees")
