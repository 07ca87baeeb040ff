"""HeatmapConfig (src/alb_sim/config/heatmap.py:4-33).  Heat maps are not part
of the traced path: run_linear discards them (execution/linear.py:117)."""
from dataclasses import dataclass

from ._meta import param


@dataclass(frozen=True)
class HeatmapConfig:
    enabled: bool = param(True, description="Whether to compute heatmaps during backward pass")
    bins: int = param(200, description="Number of bins in each dimension for the heatmaps")
    water_extent: float = param(1.0, unit="m", description="Half-width of water surface heatmap extent (centered on laser intersection)")
    seafloor_extent: float = param(10.0, unit="m", description="Half-width of seafloor heatmap extent (centered on laser intersection)")
