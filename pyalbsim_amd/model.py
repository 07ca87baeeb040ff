"""Scene model: FP64 constants derived from a SimulationConfig on the host.

Restates the constructors of the reference's physics models, i.e. everything
that is computed once per scene rather than per photon:

* laser direction          physics/models/laser.py:22-33 (angles in degrees)
* step count, time step    physics/models/simulation.py:38-74
* Fournier-Forand table    physics/scatter/fournier_forand_phase_function.py:9-48
* backscatter fraction     physics/scatter/fournier_forand_phase_function.py:89-111
* lidar attenuation, w0    physics/models/water.py:117-184, 237-245
* base reflectance F0      physics/models/sea_surface.py:30-33
* +z rotation to a vector  math/rotation_matrix.py:39-102

The traced path (HIP) and the oracle both consume the resulting ``alb_scene_t``.
Only one constant water layer is supported on the traced path; depth profiles
and layer stacks raise ``NotImplementedError`` here.
"""
from __future__ import annotations

import math
from dataclasses import dataclass

import numpy as np

from .config.simulation import SimulationConfig

LIGHT_SPEED_AIR = 2.998e8  # physics/constants.py:3
FLOAT32_EPS = float(np.finfo(np.float32).eps)  # physics/constants.py:4
SCATTER_DIVISIONS = 18000  # fournier_forand_phase_function.py:5


def laser_direction(nadir_deg: float, azimuth_deg: float) -> np.ndarray:
    """(sin t cos p, -cos t, sin t sin p) with t, p converted from degrees."""
    t = math.radians(nadir_deg)
    p = math.radians(azimuth_deg)
    return np.array([math.sin(t) * math.cos(p), -math.cos(t), math.sin(t) * math.sin(p)])


def rotation_from_z(v) -> np.ndarray:
    """float32 matrix R with R @ (0,0,1) = v / |v| (Rodrigues about z x v).

    Exactly aligned (np.isclose(c, 1)) gives I, exactly opposite gives
    diag(1, -1, -1), as rotation_matrix.py:68-77.
    """
    v = np.asarray(v, dtype=np.float64)
    n = np.linalg.norm(v)
    if n != 0:
        v = v / n
    c = v[2]
    vx, vy, vz = -v[1], v[0], 0.0
    s2 = vx * vx + vy * vy
    if np.isclose(c, 1.0):
        m = np.eye(3)
    elif np.isclose(c, -1.0):
        m = np.diag([1.0, -1.0, -1.0])
    else:
        k = np.array([[0.0, -vz, vy], [vz, 0.0, -vx], [-vy, vx, 0.0]])
        m = np.eye(3) + k + (k @ k) * ((1.0 - c) / s2)
    return m.astype(np.float32)


def fournier_forand_table(junge_slope: float, m: float):
    """(theta, p(theta), cdf) on theta = linspace(1e-6, pi, 18000).

    The CDF is the normalised running sum of p(theta) sin(theta) starting at the
    first node, so cdf[0] > 0 carries the mass of the first bin (an atom at
    theta = 1e-6 when sampling by inverse CDF).
    """
    theta = np.linspace(1e-6, np.pi, SCATTER_DIVISIONS, dtype=np.float64)
    v = (3 - junge_slope) / 2
    c = 4 / (3 * (m - 1) ** 2)
    delta = c * np.sin(theta / 2) ** 2
    delta_180 = c * np.sin(np.pi / 2) ** 2
    t1 = 1 / (4 * np.pi * (1 - delta) ** 2 * delta ** v)
    t2 = v * (1 - delta) - (1 - delta ** v) + (delta * (1 - delta ** v) - v * (1 - delta)) * np.sin(theta / 2) ** (-2)
    t3 = ((1 - delta_180 ** v) / (16 * np.pi * (delta_180 - 1) * delta_180 ** v)) * (3 * np.cos(theta) ** 2 - 1)
    pf = t1 * t2 + t3
    cdf = np.cumsum(pf * np.sin(theta))
    cdf /= cdf[-1]
    return theta, pf, cdf


def phase_bin_mass(theta: np.ndarray, cdf: np.ndarray) -> np.ndarray:
    """CDF mass of each of the len(theta) scattering-angle bins the gather uses.

    Bin i spans [i delta, (i+1) delta) with delta = float32(pi - float32 eps) /
    len(theta) in float32 arithmetic (NumPy 2 keeps pi - np.float32 in
    float32); the mass is np.interp at the upper minus np.interp at the lower
    edge (physics/scatter/energy.py:42-52), so a lookup equals that formula.
    """
    n = len(theta)
    delta = np.float32(np.float32(np.pi) - np.finfo(np.float32).eps) / np.float32(n)
    i = np.arange(n, dtype=np.int64)
    lo = i * np.float64(delta)
    hi = (i + 1) * np.float64(delta)
    return np.interp(hi, theta, cdf) - np.interp(lo, theta, cdf)


def henyey_greenstein_table(g: float):
    """(theta, p(theta), cdf) of the Henyey-Greenstein phase function on the
    Fournier-Forand grid, with the same Riemann-sum CDF, so that sampling and
    the gather's phase-mass lookup treat it exactly like the FF table.
    p = (1 - g^2) / (4 pi (1 + g^2 - 2 g cos theta)^(3/2))."""
    theta = np.linspace(1e-6, np.pi, SCATTER_DIVISIONS, dtype=np.float64)
    pf = (1 - g * g) / (4 * np.pi * (1 + g * g - 2 * g * np.cos(theta)) ** 1.5)
    cdf = np.cumsum(pf * np.sin(theta))
    cdf /= cdf[-1]
    return theta, pf, cdf


def henyey_greenstein_backscatter(g: float) -> float:
    """Backscatter fraction of HG: 2 pi int_{pi/2}^{pi} p sin = (1-g)/(2g) ((1+g)/sqrt(1+g^2) - 1)."""
    if g == 0:
        return 0.5
    return float((1 - g) / (2 * g) * ((1 + g) / math.sqrt(1 + g * g) - 1))


def cox_munk_roughness(wind_speed: float) -> float:
    """Surface roughness r (GGX alpha = r^2) from wind speed (m/s): Cox-Munk
    mean-square slope s^2 = 0.003 + 0.00512 W (clean surface, both directions);
    alpha = sqrt(s^2), the RMS slope, so r = (s^2)^(1/4)."""
    s2 = 0.003 + 0.00512 * float(wind_speed)
    return float(s2 ** 0.25)


def backscatter_fraction(junge_slope: float, m: float) -> float:
    """Closed-form FF backscatter fraction B_p."""
    v = (3 - junge_slope) / 2
    d90 = (4 / (3 * (m - 1) ** 2)) * np.sin(np.pi / 4) ** 2
    num = 1 - d90 ** (v + 1) - 0.5 * (1 - d90 ** v)
    den = (1 - d90) * d90 ** v
    return float(1 - num / den)


def lidar_attenuation(a: float, b: float, salinity: float, wavelength_um: float,
                      flying_height: float, divergence: float, bp_frac: float) -> float:
    """Churnside lidar attenuation coefficient alpha_L (1/m).

    Note the second exponential uses the total scattering coefficient b, as
    water.py:163-182 does (the authors flag it with a TODO).
    """
    bw = (1 + 0.008027 * salinity) * 0.00012 * wavelength_um ** (-4.24)
    bp = b - bw
    spot = 2 * flying_height * math.tan(divergence / 2)
    kd = a + 4.18 * bp * bp_frac * (1 - 0.52 * np.exp(-10.8 * a))
    return float(kd + (bp - 4.18 * bp * bp_frac * (1 - 0.52 * np.exp(-10.8 * b))) * np.exp(-0.85 * spot * (a + bp)))


def base_reflectance(n: float) -> float:
    return ((1 - n) / (1 + n)) ** 2


@dataclass
class SceneModel:
    """All per-scene constants; arrays stay alive as long as the model."""

    config: SimulationConfig
    y_ws: float
    y_sf: float
    depth: float
    laser_dir: np.ndarray
    rot_laser: np.ndarray  # float32 (3,3)
    divergence: float
    fov: float
    aperture_radius: float
    v_air: float
    v_water: float
    dt: float
    n_water: float
    F0: float
    roughness: float
    surface_albedo: float
    floor_albedo: float
    absorption: float
    scattering: float
    alpha_l64: float
    omega064: float
    steps: int
    mult: int
    k: int
    phase_theta: np.ndarray
    phase_pf: np.ndarray
    phase_cdf: np.ndarray
    backscatter: float
    phase_g: float | None = None  # Henyey-Greenstein asymmetry when the table is HG

    @property
    def n_substeps(self) -> int:
        return self.steps * self.mult

    @property
    def n_bins(self) -> int:
        return 2 * self.steps

    @property
    def alpha_l(self) -> float:
        """float32-rounded attenuation (the reference stores it in float32)."""
        return float(np.float32(self.alpha_l64))

    @property
    def omega0(self) -> float:
        return float(np.float32(self.omega064))

    def surface_return_bin(self) -> float:
        """Waveform bin of the nadir-ish surface return: 2 x slant range / c."""
        slant = -self.y_ws / -self.laser_dir[1]
        return 2 * slant / (self.v_air * self.dt) / self.mult

    def gate(self, length: int = 1024) -> tuple:
        """Default LDS-privatised window: starts 64 bins before the surface return."""
        g0 = max(0, int(math.floor(self.surface_return_bin())) - 64)
        g0 = min(g0, max(0, self.n_bins - length))
        return g0, min(length, self.n_bins)


def build_scene(cfg: SimulationConfig, hg_g: float | None = None) -> SceneModel:
    """Scene constants of cfg.  hg_g: use a Henyey-Greenstein phase function
    with asymmetry hg_g instead of the layer's Fournier-Forand one (its
    backscatter fraction then also enters the lidar attenuation)."""
    layers = cfg.water.layers
    if len(layers) != 1:
        raise NotImplementedError("the traced path supports exactly one water layer")
    layer = layers[0]
    for name in ("scattering_coefficient", "absorption_coefficient", "salinity", "refractive_index"):
        if not isinstance(getattr(layer, name), float):
            raise NotImplementedError(f"depth-profiled {name} is not supported on the traced path")
    depth = float(np.cumsum([layer.height])[-1])
    H = cfg.scene.flying_height
    d_l = laser_direction(cfg.laser.nadir_angle, cfg.laser.azimuth_angle)
    y_ws = -H
    y_sf = -(H + depth)
    sr = cfg.sensor.sample_rate
    mult = cfg.sample_multiplier
    dt = 1 / (mult * sr)
    distance = y_sf / d_l[1]
    steps = int(1.5 * (distance * round(sr)) / LIGHT_SPEED_AIR)
    ff = layer.fournier_forand_parameters
    if hg_g is None:
        theta, pf, cdf = fournier_forand_table(ff.junge_slope, ff.refractive_index_ratio)
        bpf = backscatter_fraction(ff.junge_slope, ff.refractive_index_ratio)
    else:
        theta, pf, cdf = henyey_greenstein_table(float(hg_g))
        bpf = henyey_greenstein_backscatter(float(hg_g))
    a = layer.absorption_coefficient
    b = layer.scattering_coefficient
    alpha = lidar_attenuation(a, b, layer.salinity, cfg.laser.wavelength, H, cfg.laser.divergence_angle, bpf)
    n = layer.refractive_index
    return SceneModel(
        config=cfg, y_ws=float(y_ws), y_sf=float(y_sf), depth=depth, laser_dir=d_l,
        rot_laser=rotation_from_z(d_l), divergence=float(cfg.laser.divergence_angle),
        fov=float(cfg.sensor.field_of_view), aperture_radius=float(cfg.sensor.aperture_radius),
        v_air=float(np.float32(LIGHT_SPEED_AIR)), v_water=float(np.float32(LIGHT_SPEED_AIR / n)),
        dt=dt, n_water=float(n), F0=base_reflectance(n), roughness=float(cfg.sea_surface.roughness),
        surface_albedo=float(cfg.sea_surface.albedo), floor_albedo=float(cfg.sea_floor.albedo),
        absorption=float(a), scattering=float(b), alpha_l64=alpha, omega064=(alpha - a) / alpha,
        steps=steps, mult=int(mult), k=int(cfg.photon_mapping_k), phase_theta=theta, phase_pf=pf,
        phase_cdf=cdf, backscatter=bpf, phase_g=None if hg_g is None else float(hg_g),
    )
