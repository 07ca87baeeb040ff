"""SimulationConfig (src/alb_sim/config/simulation.py:12-60)."""
from dataclasses import dataclass

from ._meta import param
from .heatmap import HeatmapConfig
from .laser import LaserConfig
from .scene import SceneConfig
from .sea_floor import SeaFloorConfig
from .sea_surface import SeaSurfaceConfig
from .sensor import SensorConfig
from .water import WaterConfig


@dataclass(frozen=True)
class SimulationConfig:
    sample_multiplier: int = param(10, description="Number of sub-sampling steps during sensor step")
    photon_mapping_k: int = param(100, description="Number of nearest neighbors to query during photon mapping")
    scene: SceneConfig = param(unit="SceneConfig", description="Scene config", factory=SceneConfig)
    water: WaterConfig = param(unit="WaterConfig", description="Water config", factory=WaterConfig)
    laser: LaserConfig = param(unit="LaserConfig", description="Laser config", factory=LaserConfig)
    sensor: SensorConfig = param(unit="SensorConfig", description="Sensor config", factory=SensorConfig)
    sea_floor: SeaFloorConfig = param(unit="SeaFloorConfig", description="Sea floor config", factory=SeaFloorConfig)
    sea_surface: SeaSurfaceConfig = param(unit="SeaSurfaceConfig", description="Sea surface config", factory=SeaSurfaceConfig)
    heatmap: HeatmapConfig = param(unit="HeatmapConfig", description="Sampled photon heatmap config", factory=HeatmapConfig)
