"""Scene sweeps: Latin-hypercube scene sets for the batched form.

The reference samples scene parameters with scipy's LatinHypercube in
notebooks/run_hypercube_sampling.ipynb and runs run_linear per scene.  The
sweep here draws

  depth U[1, 30] m, absorption a U[0.02, 0.5] /m, scattering b U[0.05, 2] /m,
  Henyey-Greenstein asymmetry g U[0.80, 0.95], floor albedo U[0.05, 0.4],
  wind speed U[0, 10] m/s (surface roughness via Cox-Munk)

from scipy.stats.qmc.LatinHypercube(d=6, seed=seed), in that column order, and
photons per scene 10**U[lo, hi] from numpy.random.default_rng(seed).  The
geometry is the single-shot one: 400 m flying height, nadir, 1 mrad
divergence, 20 mrad receiver half-angle, 0.1 m aperture, 0.1 ns bins,
sample multiplier 10, k = 100, water refractive index 1.34.
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .config import (HeatmapConfig, LaserConfig, SceneConfig, SeaFloorConfig, SeaSurfaceConfig, SensorConfig,
                     SimulationConfig, TurbidityLayerConfig, WaterConfig)
from .model import SceneModel, build_scene, cox_munk_roughness

RANGES = {
    "depth": (1.0, 30.0),
    "a": (0.02, 0.5),
    "b": (0.05, 2.0),
    "g": (0.80, 0.95),
    "albedo": (0.05, 0.4),
    "wind": (0.0, 10.0),
}


def shot_config(depth: float, a: float, b: float, albedo: float, roughness: float,
                sample_rate: int = int(1e10), slope: float = 0.0, azimuth: float = 0.0) -> SimulationConfig:
    return SimulationConfig(
        scene=SceneConfig(flying_height=400.0),
        laser=LaserConfig(nadir_angle=0.0, divergence_angle=1e-3, wavelength=0.532),
        sensor=SensorConfig(field_of_view=20e-3, sample_rate=int(sample_rate), aperture_radius=0.1),
        sea_surface=SeaSurfaceConfig(roughness=roughness),
        water=WaterConfig(layers=(TurbidityLayerConfig(height=depth, absorption_coefficient=a,
                                                       scattering_coefficient=b, refractive_index=1.34),)),
        sea_floor=SeaFloorConfig(albedo=albedo, slope=slope, azimuth=azimuth), sample_multiplier=10,
        photon_mapping_k=100,
        heatmap=HeatmapConfig(enabled=False))


@dataclass
class SweepScene:
    params: dict
    model: SceneModel
    n_forward: int
    n_backward: int


def hypercube(n: int = 1024, seed: int = 1, log10_photons=(5.0, 7.0), backward_ratio: float = 0.25):
    """n scenes of the sweep; n_backward = round(backward_ratio * n_forward)."""
    from scipy.stats import qmc

    u = qmc.LatinHypercube(d=6, seed=seed).random(n)
    cols = list(RANGES)
    vals = {c: RANGES[c][0] + u[:, i] * (RANGES[c][1] - RANGES[c][0]) for i, c in enumerate(cols)}
    lp = np.random.default_rng(seed).uniform(log10_photons[0], log10_photons[1], n)
    nf = np.round(10.0 ** lp).astype(np.int64)
    out = []
    for i in range(n):
        p = {c: float(vals[c][i]) for c in cols}
        p["roughness"] = cox_munk_roughness(p["wind"])
        cfg = shot_config(p["depth"], p["a"], p["b"], p["albedo"], p["roughness"])
        model = build_scene(cfg, hg_g=p["g"])
        out.append(SweepScene(p, model, int(nf[i]), int(max(1, round(backward_ratio * nf[i])))))
    return out
