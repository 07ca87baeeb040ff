"""SampledHeatmapsResult — the reference's heat-map container
(src/alb_sim/utils/heatmap_result.py:7-17), same fields."""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np


@dataclass
class SampledHeatmapsResult:
    water_heatmap: np.ndarray           # (bins, bins): gathered energy at the first water interaction
    seafloor_heatmap: np.ndarray        # (bins, bins): gathered energy at the seafloor interaction
    scatter_radius_heatmap: np.ndarray  # (bins, bins): surface radius vs seafloor radius counts
