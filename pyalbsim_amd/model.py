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
Layer stacks and depth profiles are supported (LayerModel): linear and
exponential profiles in their own arithmetic, any other profile with at(t)
through a sampled table (profile_table).
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


PROFILE_CONST, PROFILE_LINEAR, PROFILE_EXPONENTIAL, PROFILE_TABLE = 0, 1, 2, 3
PROFILE_TABLE_N = 65537  # samples of a tabulated profile over the layer's fractional depth [0, 1]


def profile_spec(value) -> tuple:
    """(kind, start, end) of a layer property: a float, a Linear/Exponential
    parameter (utils/parameter_profile.py:47-99, evaluated on the device in
    their own arithmetic), or any other ScalarParameter (an object with
    at(t), utils/parameter_profile.py:9-19), which the device interpolates
    from profile_table."""
    from .utils.parameter_profile import ExponentialParameter, LinearParameter

    if isinstance(value, float):
        return PROFILE_CONST, value, value
    if isinstance(value, LinearParameter):
        return PROFILE_LINEAR, float(value.start), float(value.end)
    if isinstance(value, ExponentialParameter):
        return PROFILE_EXPONENTIAL, float(value.start), float(value.end)
    if callable(getattr(value, "at", None)):
        tab = profile_table(value)
        return PROFILE_TABLE, float(tab[0]), float(tab[-1])
    raise TypeError(f"layer property {value!r} is neither a number nor a profile with at(t)")


def profile_table(value) -> np.ndarray:
    """FP64 samples of value.at(t) at t = i / (PROFILE_TABLE_N - 1).  Linear
    interpolation of these differs from at(t) by at most max|at''| / 8 /
    (PROFILE_TABLE_N - 1)^2 (2.9e-11 max|at''|)."""
    t = np.linspace(0.0, 1.0, PROFILE_TABLE_N)
    try:
        v = np.asarray(value.at(t), dtype=np.float64)
    except (TypeError, ValueError):
        v = None
    if v is None or v.shape != t.shape:
        v = np.array([float(value.at(float(x))) for x in t])
    if not np.all(np.isfinite(v)):
        raise ValueError(f"profile {value!r} is not finite on [0, 1]")
    return v


def eval_profile(value, z_local, height: float):
    """TurbidityLayerModel._eval (physics/models/water.py:103-109) with NumPy's
    dtype rules: floats pass through, profiles see t = z_local / height."""
    if isinstance(value, float):
        return value
    return value.at(z_local / height)


def lidar_attenuation_profile(layer, z_local, wavelength_um: float, flying_height: float, divergence: float,
                              bp_frac: float):
    """_calculate_lidar_attenuation_coefficient (physics/models/water.py:117-184)
    at local depths z_local, in the same operation order and dtypes."""
    h = layer.height
    bw = (1 + 0.008027 * eval_profile(layer.salinity, z_local, h)) * 0.00012 * wavelength_um ** (-4.24)
    bp = eval_profile(layer.scattering_coefficient, z_local, h) - bw
    spot = 2 * flying_height * math.tan(divergence / 2)
    a = eval_profile(layer.absorption_coefficient, z_local, h)
    kd = a + 4.18 * bp * bp_frac * (1 - 0.52 * np.exp(-10.8 * a))
    b = eval_profile(layer.scattering_coefficient, z_local, h)
    return kd + (bp - 4.18 * bp * bp_frac * (1 - 0.52 * np.exp(-10.8 * b))) * np.exp(-0.85 * spot * (a + bp))


@dataclass
class LayerModel:
    """One turbidity layer of a layered water column (physics/models/water.py:39-250)."""

    top: float          # depth of the top: float32 cumulative height, as layer_local_depth uses
    bottom: float       # FP64 cumulative height: depths >= bottom belong to deeper layers
    height: float
    profiled: bool
    alpha64: float      # constant layers: lidar attenuation (FP64)
    omega064: float     # constant layers: single-scattering albedo (FP64)
    v: float            # constant layers: float32(c / n)
    a_spec: tuple       # (kind, start, end) of the absorption coefficient
    n_spec: tuple       # (kind, start, end) of the refractive index
    a_tab: np.ndarray   # tabulated absorption profile (kind 3), else empty
    n_tab: np.ndarray   # tabulated refractive-index profile (kind 3), else empty
    z_lut: np.ndarray   # profiled layers: float32 z grid (num_sublayers)
    alpha_lut: np.ndarray  # profiled layers: attenuation at z_lut (as computed, float32 or float64)
    phase_theta: np.ndarray
    phase_pf: np.ndarray
    phase_cdf: np.ndarray
    backscatter: float


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
    layers: list = None           # LayerModel per layer when the column is layered or depth-profiled
    # tilted floor (extension): plane floor_normal . p = floor_offset through the beam's floor point
    tilted: bool = False
    floor_normal: np.ndarray = None
    floor_offset: float = None
    floor_t1: np.ndarray = None
    floor_t2: np.ndarray = None

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


def _layer_model(layer, top32: float, bottom: float, cfg, hg_g):
    ff = layer.fournier_forand_parameters
    if hg_g is None:
        theta, pf, cdf = fournier_forand_table(ff.junge_slope, ff.refractive_index_ratio)
        bpf = backscatter_fraction(ff.junge_slope, ff.refractive_index_ratio)
    else:
        theta, pf, cdf = henyey_greenstein_table(float(hg_g))
        bpf = henyey_greenstein_backscatter(float(hg_g))
    names = ("scattering_coefficient", "absorption_coefficient", "salinity", "refractive_index")
    profiled = not all(isinstance(getattr(layer, nm), float) for nm in names)
    for nm in names:
        profile_spec(getattr(layer, nm))  # rejects values that are neither numbers nor profiles
    H = cfg.scene.flying_height
    lam, div = cfg.laser.wavelength, cfg.laser.divergence_angle
    a_spec, n_spec = profile_spec(layer.absorption_coefficient), profile_spec(layer.refractive_index)
    if profiled:
        if int(layer.num_sublayers) < 1:
            raise ValueError("a depth-profiled layer needs num_sublayers >= 1")
        z_lut = np.linspace(0.0, layer.height, int(layer.num_sublayers), dtype=np.float32)
        # the reference's backscatter fraction is a NumPy float64 scalar, which takes part in dtype promotion
        lut = np.asarray(lidar_attenuation_profile(layer, z_lut, lam, H, div, np.float64(bpf)))
        if lut.ndim == 0:
            lut = np.full(z_lut.shape, lut)
        alpha64 = omega064 = v = float("nan")
    else:
        z_lut = np.zeros(0, dtype=np.float32)
        lut = np.zeros(0)
        a = layer.absorption_coefficient
        alpha64 = lidar_attenuation(a, layer.scattering_coefficient, layer.salinity, lam, H, div, bpf)
        omega064 = (alpha64 - a) / alpha64
        v = float(np.float32(LIGHT_SPEED_AIR / layer.refractive_index))
    return LayerModel(top=float(top32), bottom=float(bottom), height=float(layer.height), profiled=profiled,
                      alpha64=float(alpha64), omega064=float(omega064), v=v, a_spec=a_spec, n_spec=n_spec,
                      a_tab=profile_table(layer.absorption_coefficient) if a_spec[0] == PROFILE_TABLE
                      else np.zeros(0),
                      n_tab=profile_table(layer.refractive_index) if n_spec[0] == PROFILE_TABLE else np.zeros(0),
                      z_lut=z_lut, alpha_lut=lut, phase_theta=theta, phase_pf=pf, phase_cdf=cdf, backscatter=bpf)


def layer_arrays(model: "SceneModel") -> list:
    """Contiguous per-layer tables of a layered scene, in the dtypes alb_layer_t
    points at: z grid (float32), attenuation at the grid (FP64), phase theta,
    CDF and phase-bin mass (FP64)."""
    out = []
    for lm in model.layers or ():
        out.append({
            "z_lut": np.ascontiguousarray(lm.z_lut, dtype=np.float32),
            "alpha_lut": np.ascontiguousarray(lm.alpha_lut, dtype=np.float64),
            "theta": np.ascontiguousarray(lm.phase_theta, dtype=np.float64),
            "cdf": np.ascontiguousarray(lm.phase_cdf, dtype=np.float64),
            "mass": np.ascontiguousarray(phase_bin_mass(lm.phase_theta, lm.phase_cdf), dtype=np.float64),
            # at least one entry each, so that every table has an address
            "a_tab": np.ascontiguousarray(lm.a_tab if len(lm.a_tab) else np.zeros(1), dtype=np.float64),
            "n_tab": np.ascontiguousarray(lm.n_tab if len(lm.n_tab) else np.zeros(1), dtype=np.float64),
        })
    return out


def beam_floor_point(d_l: np.ndarray, y_ws: float, depth: float, n: float) -> np.ndarray:
    """Point where the laser boresight, refracted at the surface with eta = 1/n
    (physics/reflection/snell_refraction.py:7-45), meets the flat floor
    (SimulationModel.sea_floor_intersection, physics/models/simulation.py:104-124)."""
    d = np.asarray(d_l, dtype=np.float64)
    eta = 1.0 / n
    normal = np.array([0.0, 1.0, 0.0])
    cos_i = float(np.clip(np.dot(-d, normal), -1.0, 1.0))
    sin2_t = eta ** 2 * (1.0 - cos_i ** 2)
    cos_t = np.sqrt(max(0.0, 1.0 - sin2_t))
    t = eta * d + (eta * cos_i - cos_t) * normal
    t = t / np.linalg.norm(t)
    surface = (y_ws / d[1]) * d
    return surface + (-depth / t[1]) * t


def floor_plane(slope_deg: float, azimuth_deg: float, point: np.ndarray):
    """(normal, offset, t1, t2) of a floor plane through point, tilted by
    slope about the horizontal, its normal leaning toward azimuth in the x-z
    plane; (t1, normal, t2) is orthonormal and equals the axes for slope 0,
    azimuth 0."""
    s, a = math.radians(slope_deg), math.radians(azimuth_deg)
    n = np.array([math.sin(s) * math.cos(a), math.cos(s), math.sin(s) * math.sin(a)])
    t1 = np.array([math.cos(s) * math.cos(a), -math.sin(s), math.cos(s) * math.sin(a)])
    t2 = np.array([-math.sin(a), 0.0, math.cos(a)])
    off = float(n[0] * point[0] + n[1] * point[1] + n[2] * point[2])
    return n, off, t1, t2


def build_scene(cfg: SimulationConfig, hg_g: float | None = None) -> SceneModel:
    """Scene constants of cfg.  hg_g: use a Henyey-Greenstein phase function
    with asymmetry hg_g instead of the layers' Fournier-Forand ones (its
    backscatter fraction then also enters the lidar attenuation).

    A column of one constant layer uses the closed-form water legs; several
    layers or depth profiles (physics/models/water.py:253-416) are traced one
    sub-step at a time with the per-depth lookups (SceneModel.layers)."""
    layers = cfg.water.layers
    heights = [layer.height for layer in layers]
    cum = np.cumsum(heights)
    depth = float(cum[-1])
    tops32 = np.concatenate(([0.0], cum[:-1])).astype(np.float32)
    lmodels = [_layer_model(layer, tops32[i], cum[i], cfg, hg_g) for i, layer in enumerate(layers)]
    slope = float(getattr(cfg.sea_floor, "slope", 0.0))
    tilted = slope > 0.0
    layered = len(layers) > 1 or lmodels[0].profiled
    layer = layers[0]
    l0 = lmodels[0]
    H = cfg.scene.flying_height
    d_l = laser_direction(cfg.laser.nadir_angle, cfg.laser.azimuth_angle)
    y_ws = -H
    y_sf = -(H + depth)
    sr = cfg.sensor.sample_rate
    mult = cfg.sample_multiplier
    dt = 1 / (mult * sr)
    distance = y_sf / d_l[1]
    steps = int(1.5 * (distance * round(sr)) / LIGHT_SPEED_AIR)
    # sea surface: layer 0's refractive index at z = 0 (physics/models/sea_surface.py:30-33)
    n = float(eval_profile(layer.refractive_index, 0.0, layer.height))
    a0 = float(eval_profile(layer.absorption_coefficient, 0.0, layer.height))
    b0 = float(eval_profile(layer.scattering_coefficient, 0.0, layer.height))
    alpha = l0.alpha64 if not l0.profiled else float("nan")
    return SceneModel(
        config=cfg, y_ws=float(y_ws), y_sf=float(y_sf), depth=depth, laser_dir=d_l,
        rot_laser=rotation_from_z(d_l), divergence=float(cfg.laser.divergence_angle),
        fov=float(cfg.sensor.field_of_view), aperture_radius=float(cfg.sensor.aperture_radius),
        v_air=float(np.float32(LIGHT_SPEED_AIR)), v_water=l0.v if not l0.profiled else float("nan"),
        dt=dt, n_water=n, F0=base_reflectance(n), roughness=float(cfg.sea_surface.roughness),
        surface_albedo=float(cfg.sea_surface.albedo), floor_albedo=float(cfg.sea_floor.albedo),
        absorption=a0, scattering=b0, alpha_l64=alpha, omega064=(alpha - a0) / alpha if not l0.profiled else float("nan"),
        steps=steps, mult=int(mult), k=int(cfg.photon_mapping_k), phase_theta=l0.phase_theta, phase_pf=l0.phase_pf,
        phase_cdf=l0.phase_cdf, backscatter=l0.backscatter, phase_g=None if hg_g is None else float(hg_g),
        layers=lmodels if layered else None,
        **_floor_fields(tilted, slope, float(getattr(cfg.sea_floor, "azimuth", 0.0)), d_l, y_ws, y_sf, depth, n),
    )


def _floor_fields(tilted, slope, azimuth, d_l, y_ws, y_sf, depth, n) -> dict:
    if not tilted:
        return dict(tilted=False, floor_normal=np.array([0.0, 1.0, 0.0]), floor_offset=float(y_sf),
                    floor_t1=np.array([1.0, 0.0, 0.0]), floor_t2=np.array([0.0, 0.0, 1.0]))
    nrm, off, t1, t2 = floor_plane(slope, azimuth, beam_floor_point(d_l, y_ws, depth, n))
    return dict(tilted=True, floor_normal=nrm, floor_offset=off, floor_t1=t1, floor_t2=t2)
