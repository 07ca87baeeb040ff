"""Scene configurations used by the tests (sizes from the benchmark workloads)."""
import numpy as np

from pyalbsim_amd.config import (HeatmapConfig, LaserConfig, SceneConfig, SeaFloorConfig, SeaSurfaceConfig,
                                 SensorConfig, SimulationConfig, TurbidityLayerConfig, WaterConfig)


def single_shot(flying_height=400.0, depth=5.0, a=0.1, b=0.2, albedo=0.15, roughness=0.05, nadir=0.0,
                fov=20e-3, sample_rate=int(1e10), k=100, n_water=1.34):
    """Plain single ALB shot: 400 m, nadir, 1 mrad, 20 mrad FOV, 5 m water,
    a = 0.1/m, b = 0.2/m, Lambertian floor albedo 0.15, 0.1 ns bins."""
    return SimulationConfig(
        scene=SceneConfig(flying_height=flying_height),
        laser=LaserConfig(nadir_angle=nadir, divergence_angle=1e-3, wavelength=0.532),
        sensor=SensorConfig(field_of_view=fov, sample_rate=sample_rate, aperture_radius=0.1),
        sea_surface=SeaSurfaceConfig(roughness=roughness),
        water=WaterConfig(layers=(TurbidityLayerConfig(height=depth, absorption_coefficient=a,
                                                       scattering_coefficient=b, refractive_index=n_water),)),
        sea_floor=SeaFloorConfig(albedo=albedo), sample_multiplier=10, photon_mapping_k=k,
        heatmap=HeatmapConfig(enabled=False))


def run_linear_script_scene():
    """The scene of scripts/run_linear.py:12-21 (reference defaults otherwise)."""
    return SimulationConfig(
        water=WaterConfig(layers=(TurbidityLayerConfig(height=7, absorption_coefficient=0.01,
                                                       scattering_coefficient=3.5),)),
        sea_floor=SeaFloorConfig(albedo=0.1))


class SmoothStepParameter:
    """A depth profile that is neither linear nor exponential: start + (end -
    start) (3 t^2 - 2 t^3), with at(t) like the reference's ScalarParameter
    protocol (utils/parameter_profile.py:9-19)."""

    def __init__(self, start: float, end: float):
        self.start, self.end = float(start), float(end)

    def at(self, t):
        t = np.asarray(t, dtype=np.float64)
        v = self.start + (self.end - self.start) * (3.0 * t * t - 2.0 * t * t * t)
        return float(v) if v.ndim == 0 else v


def layered_scene(n_profile=None, k=100, flying_height=400.0, nadir=0.0, a_profile=None, b_profile=None):
    """Three-layer column, 6 m: a clear constant top layer, a depth-profiled
    middle layer (linear scattering, exponential absorption, linear refractive
    index unless n_profile is given) with its own Fournier-Forand particles,
    and a turbid constant bottom layer."""
    from pyalbsim_amd.config import FournierForandConfig
    from pyalbsim_amd.utils.parameter_profile import ExponentialParameter, LinearParameter

    n_mid = LinearParameter(1.34, 1.345) if n_profile is None else n_profile
    layers = (
        TurbidityLayerConfig(height=1.5, absorption_coefficient=0.08, scattering_coefficient=0.25,
                             refractive_index=1.34),
        TurbidityLayerConfig(height=2.0,
                             scattering_coefficient=LinearParameter(0.3, 1.2) if b_profile is None else b_profile,
                             absorption_coefficient=ExponentialParameter(0.05, 0.2) if a_profile is None else a_profile,
                             salinity=35.0,
                             refractive_index=n_mid, num_sublayers=40,
                             fournier_forand_parameters=FournierForandConfig(junge_slope=3.8,
                                                                             refractive_index_ratio=1.05)),
        TurbidityLayerConfig(height=2.5, absorption_coefficient=0.15, scattering_coefficient=0.6,
                             refractive_index=1.34,
                             fournier_forand_parameters=FournierForandConfig(junge_slope=3.4,
                                                                             refractive_index_ratio=1.15)),
    )
    return SimulationConfig(
        scene=SceneConfig(flying_height=flying_height),
        laser=LaserConfig(nadir_angle=nadir, divergence_angle=1e-3, wavelength=0.532),
        sensor=SensorConfig(field_of_view=20e-3, sample_rate=int(1e10), aperture_radius=0.1),
        sea_surface=SeaSurfaceConfig(roughness=0.05),
        water=WaterConfig(layers=layers),
        sea_floor=SeaFloorConfig(albedo=0.15), sample_multiplier=10, photon_mapping_k=k,
        heatmap=HeatmapConfig(enabled=False))


def stacked_copies(cfg, n: int):
    """cfg with its single constant layer split into n identical layers of
    the same total height."""
    import dataclasses

    (layer,) = cfg.water.layers
    part = dataclasses.replace(layer, height=layer.height / n)
    return dataclasses.replace(cfg, water=WaterConfig(layers=(part,) * n))


def tilted_shot(slope=10.0, azimuth=40.0, **kw):
    """single_shot with the floor tilted by slope degrees toward azimuth (extension)."""
    import dataclasses

    cfg = single_shot(**kw)
    return dataclasses.replace(cfg, sea_floor=SeaFloorConfig(albedo=cfg.sea_floor.albedo, slope=slope,
                                                             azimuth=azimuth))
