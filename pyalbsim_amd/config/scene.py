"""SceneConfig (src/alb_sim/config/scene.py:4-16)."""
from dataclasses import dataclass

from ._meta import param


@dataclass(frozen=True)
class SceneConfig:
    flying_height: float = param(135, unit="m", description="Height above water of sensor and laser")

    def __post_init__(self):
        if self.flying_height <= 0:
            raise ValueError("Flying height must be positive")
