"""Photon interaction categories, in the reference's enum order.

src/alb_sim/photon_mapping/photon_type.py:4-10: BOTTOM_REFLECTION, SCATTER,
SURFACE_REFLECTION, SURFACE_TRANSMISSION_UP (auto() values 1..4).  The C ABI
indexes the same members 0..3 (``PhotonType.X.value - 1``).
"""
from enum import Enum, auto


class PhotonType(Enum):
    BOTTOM_REFLECTION = auto()
    SCATTER = auto()
    SURFACE_REFLECTION = auto()
    SURFACE_TRANSMISSION_UP = auto()

    @property
    def abi_index(self) -> int:
        return self.value - 1
