"""Frozen configuration dataclasses: same class names, field names, defaults,
units and validation as the reference's src/alb_sim/config/*.py, so scripts
written against ``alb_sim.config`` construct the same scenes."""
from .heatmap import HeatmapConfig
from .laser import LaserConfig
from .run import RunConfig
from .scene import SceneConfig
from .sea_floor import SeaFloorConfig
from .sea_surface import SeaSurfaceConfig
from .sensor import SensorConfig
from .simulation import SimulationConfig
from .water import FournierForandConfig, TurbidityLayerConfig, WaterConfig

__all__ = [
    "HeatmapConfig", "LaserConfig", "RunConfig", "SceneConfig", "SeaFloorConfig",
    "SeaSurfaceConfig", "SensorConfig", "SimulationConfig", "FournierForandConfig",
    "TurbidityLayerConfig", "WaterConfig",
]
