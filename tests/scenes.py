"""Scene configurations used by the tests (sizes from the benchmark workloads)."""
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
