"""ctypes mirror of include/alb_abi.h (layout checked against alb_scene_size())."""
from __future__ import annotations

import ctypes as C

import numpy as np

ALB_ABI_VERSION = 11
KNN_WORK_WORDS = 256  # ALB_KNN_WORK_WORDS: scratch words of alb_knn / alb_knn_batched
GATHER_STATIC_MAX_QUERIES = 1 << 20  # ALB_GATHER_STATIC_MAX_QUERIES: queries per alb_gather_static call
MAX_K = 960  # largest photon_mapping_k the k-NN kernels take (LDS candidate buffer of 1024)
GATHER_STATIC_SCRATCH_BYTES = 64 << 20  # ALB_GATHER_STATIC_SCRATCH_BYTES: its static device array per GPU
N_TYPES = 4
N_COUNTERS = 19

ALB_OK = 0
ALB_ERR_ARG = -1
ALB_ERR_CAPACITY = -2
ALB_ERR_HIP = -3
ALB_ERR_UNSUPPORTED = -4
ALB_ERR_WORKSPACE = -5

BOTTOM_REFLECTION, SCATTER, SURFACE_REFLECTION, SURFACE_TRANSMISSION_UP = range(4)
INTERACTION_FIELDS = ("fwi_x", "fwi_y", "fwi_z", "sfi_x", "sfi_y", "sfi_z")
TYPE_NAMES = ("BOTTOM_REFLECTION", "SCATTER", "SURFACE_REFLECTION", "SURFACE_TRANSMISSION_UP")

RNG_PHILOX = 0
RNG_REPLAY = 1
PASS_FORWARD = 0
PASS_BACKWARD = 1

CTR_PATHS, CTR_AIR_LEGS, CTR_WATER_STEPS, CTR_SCATTER, CTR_ENTER, CTR_EXIT, CTR_FLOOR, \
    CTR_REPLAY_EXHAUSTED, CTR_QUERIES, CTR_CANDIDATES, CTR_NODES, CTR_OVERFLOW, CTR_DEPTH_SOLVES, \
    CTR_KNN_STEPS, CTR_KNN_SORTS, CTR_KNN_DESCENT, CTR_SMALL_SCENE, CTR_REPLAY_MISSES, CTR_NEIGHBOURS = range(19)
COUNTER_NAMES = ["paths", "air_legs", "water_steps", "scatter", "enter", "exit", "floor",
                 "replay_exhausted", "queries", "candidates", "nodes", "overflow", "depth_solves",
                 "knn_steps", "knn_sorts", "knn_descent", "small_scene", "replay_misses", "neighbours"]


class Scene(C.Structure):
    _fields_ = [
        ("y_ws", C.c_double), ("y_sf", C.c_double), ("laser_dir", C.c_double * 3),
        ("rot_laser", C.c_float * 9), ("pad0", C.c_float),
        ("divergence", C.c_double), ("rayleigh_sigma", C.c_double), ("cos_fov", C.c_double),
        ("aperture_radius", C.c_double), ("v_air", C.c_double), ("v_water", C.c_double),
        ("dt", C.c_double), ("n_water", C.c_double), ("eta", C.c_double), ("F0", C.c_double),
        ("alpha_ggx", C.c_double), ("surface_albedo", C.c_double), ("floor_albedo", C.c_double),
        ("absorption", C.c_double), ("alpha_l", C.c_float), ("omega0", C.c_float),
        ("n_substeps", C.c_int64), ("steps", C.c_int32), ("mult", C.c_int32), ("k", C.c_int32),
        ("phase_n", C.c_int32), ("phase_theta", C.c_void_p), ("phase_cdf", C.c_void_p),
        ("phase_mass", C.c_void_p), ("phase_score", C.c_void_p), ("phase_dlog_mass", C.c_void_p),
        ("n_layers", C.c_int32), ("pad1", C.c_int32), ("layers", C.c_void_p),
        ("heat_bins", C.c_int32), ("pad2", C.c_int32), ("heat_water_extent", C.c_double),
        ("heat_floor_extent", C.c_double), ("heat_water_center", C.c_double * 3),
        ("heat_floor_center", C.c_double * 3), ("heat_radius", C.c_void_p), ("heat_water", C.c_void_p),
        ("heat_floor", C.c_void_p), ("tilted", C.c_int32), ("pad3", C.c_int32),
        ("floor_normal", C.c_double * 3), ("floor_offset", C.c_double), ("floor_t1", C.c_double * 3),
        ("floor_t2", C.c_double * 3),
    ]


class Layer(C.Structure):
    _fields_ = [
        ("top", C.c_double), ("bottom", C.c_double), ("height", C.c_double), ("profiled", C.c_int32),
        ("lut_n", C.c_int32), ("alpha", C.c_float), ("omega0", C.c_float), ("v", C.c_double),
        ("a_kind", C.c_int32), ("n_kind", C.c_int32), ("a0", C.c_double), ("a1", C.c_double),
        ("n0", C.c_double), ("n1", C.c_double), ("z_lut", C.c_void_p), ("alpha_lut", C.c_void_p),
        ("phase_theta", C.c_void_p), ("phase_cdf", C.c_void_p), ("phase_mass", C.c_void_p),
        ("phase_n", C.c_int32), ("pad0", C.c_int32), ("a_tab", C.c_void_p), ("n_tab", C.c_void_p),
        ("tab_n", C.c_int32), ("pad1", C.c_int32),
    ]


def make_layer(lm, z_lut_ptr: int, alpha_lut_ptr: int, theta_ptr: int, cdf_ptr: int, mass_ptr: int,
               a_tab_ptr: int = 0, n_tab_ptr: int = 0) -> Layer:
    """Pack a model.LayerModel; the table pointers are device or host addresses."""
    L = Layer()
    L.top, L.bottom, L.height = lm.top, lm.bottom, lm.height
    L.profiled = 1 if lm.profiled else 0
    L.lut_n = len(lm.z_lut)
    L.alpha = float(np.float32(lm.alpha64)) if not lm.profiled else 0.0
    L.omega0 = float(np.float32(lm.omega064)) if not lm.profiled else 0.0
    L.v = lm.v if not lm.profiled else 0.0
    L.a_kind, L.a0, L.a1 = lm.a_spec
    L.n_kind, L.n0, L.n1 = lm.n_spec
    L.z_lut, L.alpha_lut = z_lut_ptr, alpha_lut_ptr
    L.phase_theta, L.phase_cdf, L.phase_mass = theta_ptr, cdf_ptr, mass_ptr
    L.phase_n = len(lm.phase_theta)
    L.a_tab = a_tab_ptr if len(lm.a_tab) else 0
    L.n_tab = n_tab_ptr if len(lm.n_tab) else 0
    L.tab_n = max(len(lm.a_tab), len(lm.n_tab))
    return L


LAYER_TABLES = ("z_lut", "alpha_lut", "theta", "cdf", "mass", "a_tab", "n_tab")


def make_layers(model, ptrs) -> C.Array:
    """alb_layer_t[n_layers] of a layered SceneModel; ptrs[i] maps the names in
    LAYER_TABLES to layer i's table addresses (device or host)."""
    arr = (Layer * len(model.layers))()
    for i, lm in enumerate(model.layers):
        p = ptrs[i]
        arr[i] = make_layer(lm, p["z_lut"], p["alpha_lut"], p["theta"], p["cdf"], p["mass"], p["a_tab"], p["n_tab"])
    return arr


class Rng(C.Structure):
    _fields_ = [("mode", C.c_int32), ("key", C.c_uint32 * 2), ("stream", C.c_uint32),
                ("replay", C.c_void_p), ("replay_len", C.c_int64)]


class Records(C.Structure):
    _fields_ = [("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("dx", C.c_void_p),
                ("dy", C.c_void_p), ("dz", C.c_void_p), ("energy", C.c_void_p), ("time", C.c_void_p),
                ("tag", C.c_void_p), ("capacity", C.c_int64), ("count", C.c_void_p), ("scene", C.c_void_p)] + \
        [(f, C.c_void_p) for f in ("fwi_x", "fwi_y", "fwi_z", "sfi_x", "sfi_y", "sfi_z")]


class Queries(C.Structure):
    _fields_ = list(Records._fields_)


class Map(C.Structure):
    _fields_ = [("n", C.c_int64), ("dims", C.c_int32), ("pad0", C.c_int32),
                ("x", C.c_void_p), ("y", C.c_void_p), ("z", C.c_void_p), ("dx", C.c_void_p),
                ("dy", C.c_void_p), ("dz", C.c_void_p), ("energy", C.c_void_p), ("time", C.c_void_p),
                ("perm", C.c_void_p), ("keys", C.c_void_p), ("meta", C.c_void_p),
                ("nodes", C.c_void_p)] + [(f, C.c_void_p) for f in ("fwi_x", "fwi_y", "fwi_z", "sfi_x", "sfi_y",
                                                                     "sfi_z")]


class SceneIndex(C.Structure):
    _fields_ = [("n_scenes", C.c_int32), ("scene_bits", C.c_int32), ("first", C.c_void_p), ("last", C.c_void_p),
                ("root", C.c_void_p)]


class ReplayIndex(C.Structure):
    _fields_ = [("keys", C.c_void_p), ("slots", C.c_void_p), ("n", C.c_int64), ("weight", C.c_void_p)]


def make_scene(model, theta_ptr: int, cdf_ptr: int, mass_ptr: int = 0, score_ptr: int = 0,
               dlog_mass_ptr: int = 0) -> Scene:
    """Pack a SceneModel; the phase-table pointers are device or host addresses."""
    s = Scene()
    s.y_ws = model.y_ws
    s.y_sf = model.y_sf
    for i in range(3):
        s.laser_dir[i] = float(model.laser_dir[i])
    rot = np.asarray(model.rot_laser, dtype=np.float32).reshape(9)
    for i in range(9):
        s.rot_laser[i] = float(rot[i])
    s.divergence = model.divergence
    s.rayleigh_sigma = model.divergence / 3.0
    s.cos_fov = float(np.cos(model.fov))
    s.aperture_radius = model.aperture_radius
    s.v_air = model.v_air
    s.v_water = model.v_water
    s.dt = model.dt
    s.n_water = model.n_water
    s.eta = 1.0 / model.n_water
    s.F0 = model.F0
    s.alpha_ggx = model.roughness * model.roughness
    s.surface_albedo = model.surface_albedo
    s.floor_albedo = model.floor_albedo
    s.absorption = model.absorption
    s.alpha_l = model.alpha_l
    s.omega0 = model.omega0
    s.n_substeps = model.n_substeps
    s.steps = model.steps
    s.mult = model.mult
    s.k = model.k
    s.phase_n = len(model.phase_theta)
    s.phase_theta = theta_ptr
    s.phase_cdf = cdf_ptr
    s.phase_mass = mass_ptr
    s.phase_score = score_ptr
    s.phase_dlog_mass = dlog_mass_ptr
    s.tilted = 1 if getattr(model, "tilted", False) else 0
    fn = getattr(model, "floor_normal", None)
    fn = (0.0, 1.0, 0.0) if fn is None else fn
    t1 = getattr(model, "floor_t1", None)
    t1 = (1.0, 0.0, 0.0) if t1 is None else t1
    t2 = getattr(model, "floor_t2", None)
    t2 = (0.0, 0.0, 1.0) if t2 is None else t2
    for i in range(3):
        s.floor_normal[i] = float(fn[i])
        s.floor_t1[i] = float(t1[i])
        s.floor_t2[i] = float(t2[i])
    s.floor_offset = float(getattr(model, "floor_offset", model.y_sf))
    return s


def make_rng(key, stream: int = 0, replay_ptr: int = 0, replay_len: int = 0) -> Rng:
    r = Rng()
    if replay_ptr:
        r.mode = RNG_REPLAY
        r.replay = replay_ptr
        r.replay_len = int(replay_len)
    else:
        r.mode = RNG_PHILOX
    r.key[0] = int(key[0]) & 0xFFFFFFFF
    r.key[1] = int(key[1]) & 0xFFFFFFFF
    r.stream = int(stream) & 0xFFFFFFFF
    return r


def key_from_generator(rng: np.random.Generator):
    """Philox key derived from a numpy Generator (consumes two draws)."""
    a = int(rng.integers(0, 2 ** 32))
    b = int(rng.integers(0, 2 ** 32))
    return (a, b)
