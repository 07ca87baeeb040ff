"""Dense float32 SoA of one photon map (photon_mapping/photon_map_data.py:6-15).

``first_water_interaction`` and ``seafloor_interaction`` feed only the heat
maps, which ``run_linear`` discards; maps built on the GPU leave them NaN.
"""
from dataclasses import dataclass

import numpy as np


@dataclass
class PhotonMapData:
    positions: np.ndarray  # (N, 3) float32
    directions: np.ndarray  # (N, 3) float32
    energies: np.ndarray  # (N,) float32
    times: np.ndarray  # (N,) float32, sub-step index
    first_water_interaction: np.ndarray  # (N, 3) float32
    seafloor_interaction: np.ndarray  # (N, 3) float32

    def __len__(self) -> int:
        return int(self.positions.shape[0])
