"""Dense float32 SoA of one photon map (photon_mapping/photon_map_data.py:6-15).

``first_water_interaction`` and ``seafloor_interaction`` are each record's
photon's first water interaction (its surface entry point) and latest
seafloor interaction, NaN before the interaction; they feed only the heat
maps.  ``run_parallel`` fills them from the GPU trace.
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


_FIELDS = ("positions", "directions", "energies", "times", "first_water_interaction", "seafloor_interaction")


def save_photon_maps(path: str, maps: dict) -> None:
    """Write {PhotonType: PhotonMapData} as one .npz of float32 arrays named
    "<TYPE NAME>/<field>" (no pickled objects)."""
    arrays = {}
    for pt, m in maps.items():
        for f in _FIELDS:
            arrays[f"{pt.name}/{f}"] = np.ascontiguousarray(getattr(m, f), dtype=np.float32)
    np.savez(path, **arrays)


def load_photon_maps(path: str) -> dict:
    """Inverse of save_photon_maps (loaded with allow_pickle=False)."""
    from .photon_type import PhotonType

    with np.load(path, allow_pickle=False) as z:
        names = {k.split("/")[0] for k in z.files}
        return {PhotonType[n]: PhotonMapData(**{f: z[f"{n}/{f}"] for f in _FIELDS}) for n in sorted(names)}
