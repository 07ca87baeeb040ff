"""Footprint-grid queries for the fixed-radius gather.

A grid x grid set of query points on the seafloor plane (the tilted plane
n . p = c where the floor is tilted, over the same (x, z) grid), centred on the
laser's seafloor intersection (SimulationModel.sea_floor_intersection,
src/alb_sim/physics/models/simulation.py:104-124: the boresight refracted at
the surface with eta = 1 / n of layer 0, physics/reflection/snell_refraction.py:7-45),
with the given pitch.  Each query looks down from the sensor at the origin
(direction = unit vector from the origin to the point), carries energy 1 and
the ballistic time of that straight path in sub-steps: (air length / v_air +
water length / v_water) / dt, with v_water the height-weighted harmonic mean
of the layers' speeds at their mid-depths for a layered column.
"""
from __future__ import annotations

import numpy as np

from .model import LIGHT_SPEED_AIR, SceneModel, beam_floor_point, eval_profile


def water_surface_intersection(model: SceneModel) -> np.ndarray:
    """Point where the laser boresight meets the water surface
    (SimulationModel.water_surface_intersection, physics/models/simulation.py:90-101)."""
    d = np.asarray(model.laser_dir, dtype=np.float64)
    return (model.y_ws / d[1]) * d


def seafloor_intersection(model: SceneModel) -> np.ndarray:
    """Point where the laser boresight, refracted at the surface, meets the floor."""
    return beam_floor_point(model.laser_dir, model.y_ws, model.depth, model.n_water)


def _water_speed(model: SceneModel) -> float:
    if not model.layers:
        return model.v_water
    cfg_layers = model.config.water.layers
    inv = 0.0
    for layer in cfg_layers:
        n = float(eval_profile(layer.refractive_index, 0.5 * layer.height, layer.height))
        inv += layer.height * n / LIGHT_SPEED_AIR
    return model.depth / inv


def footprint_queries(model: SceneModel, grid: int = 256, pitch: float = 0.01) -> dict:
    """Host arrays of the footprint queries (x, y, z, dx, dy, dz FP64; energy,
    time float32), row-major over (x, z)."""
    c = seafloor_intersection(model)
    off = (np.arange(grid, dtype=np.float64) - (grid - 1) / 2.0) * pitch
    gx, gz = np.meshgrid(c[0] + off, c[2] + off, indexing="ij")
    x = gx.ravel()
    z = gz.ravel()
    if getattr(model, "tilted", False):
        nrm = np.asarray(model.floor_normal, dtype=np.float64)
        y = (model.floor_offset - nrm[0] * x - nrm[2] * z) / nrm[1]
    else:
        y = np.full_like(x, model.y_sf)
    length = np.sqrt(x * x + y * y + z * z)
    dx, dy, dz = x / length, y / length, z / length
    air = length * (model.y_ws / y)  # the straight sensor -> point path crosses the surface at y_ws
    water = length - air
    steps = (air / model.v_air + water / _water_speed(model)) / model.dt
    return {"x": x, "y": y, "z": z, "dx": dx, "dy": dy, "dz": dz,
            "energy": np.ones_like(x, dtype=np.float32), "time": steps.astype(np.float32)}


def upload_queries(arrs: dict, device="cuda"):
    """A device query sink holding host query arrays (runtime.Sink)."""
    import torch

    from . import runtime

    n = len(arrs["x"])
    sink = runtime.Sink.alloc(n, query=True, tags=False, device=device)
    for f in runtime.REC_F32:
        sink.tensors[f][:n].copy_(torch.as_tensor(np.ascontiguousarray(arrs[f])))
    sink.tensors["count"].fill_(n)
    return sink
