"""CPU oracle for the photon-tracing / photon-mapping path — TEST INFRASTRUCTURE.

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
bench.py may import this package, and only as the checker (or as the timed CPU
restatement).  The product (pyalbsim_amd) never imports it.

The numerics live in alb_oracle.c (a literal, time-stepped restatement of
src/alb_sim/core/simulation.py); this module builds it with gcc, loads it with
ctypes and wraps it in NumPy-friendly calls.  Function-by-function citations
and pinning are in alb_oracle.c and tests/test_oracle_*.py.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess
import threading

import numpy as np

from pyalbsim_amd import abi

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libalb_oracle.so")
_lib = None
_lock = threading.Lock()


def build(force: bool = False) -> str:
    """Compile alb_oracle.c with gcc (oracle/Makefile).  A file lock keeps
    concurrent processes (parallel test workers) from building at once."""
    import fcntl

    srcs = (os.path.join(HERE, "alb_oracle.c"), os.path.join(HERE, "..", "include", "alb_abi.h"))

    def stale():
        return not os.path.exists(LIB_PATH) or os.path.getmtime(LIB_PATH) < max(os.path.getmtime(p) for p in srcs)

    if force or stale():
        os.makedirs(os.path.dirname(LIB_PATH), exist_ok=True)
        with open(LIB_PATH + ".lock", "w") as lk:
            fcntl.flock(lk, fcntl.LOCK_EX)
            if force or stale():
                subprocess.run(["make", "-s", "-C", HERE, "-B" if force else "all"], check=True)
    return LIB_PATH


def lib():
    global _lib
    with _lock:
        if _lib is None:
            build()
            L = C.CDLL(LIB_PATH)
            P = C.c_void_p
            L.orc_philox4x32_10.argtypes = [P, P, P]
            L.orc_uniform_at.argtypes = [P, C.c_uint64, C.c_uint64]
            L.orc_uniform_at.restype = C.c_double
            L.orc_trace.argtypes = [P, P, C.c_int32, C.c_int64, C.c_int64, P, P, P]
            L.orc_trace.restype = C.c_int64
            L.orc_trace_stats.argtypes = [P, P, C.c_int32, C.c_int64, C.c_int64, P, P, P, P, C.c_int32]
            L.orc_trace_stats.restype = C.c_int64
            L.orc_knn_one.argtypes = [P, P, P, C.c_int64, C.c_double, C.c_double, C.c_double,
                                      C.c_int32, P, P]
            L.orc_knn_one.restype = C.c_int
            L.orc_gather.argtypes = [P, C.c_int32, P, P, P, P, P, P, P, P, C.c_int64, P,
                                     C.c_int64, P, C.c_int64, P, P]
            L.orc_gather.restype = C.c_int
            L.orc_gather_ex.argtypes = [P, C.c_int32, P, P, P, P, P, P, P, P, C.c_int64, P,
                                        C.c_int64, P, C.c_int64, P, P, P, P, P, P, P]
            L.orc_gather_ex.restype = C.c_int
            L.orc_radius_gather.argtypes = [P, C.c_int32, P, P, P, P, P, P, P, P, C.c_int64, P,
                                            C.c_int64, C.c_double, P, C.c_int64, P, P]
            L.orc_radius_gather.restype = C.c_int
            L.orc_rotation_from_z.argtypes = [P, P]
            L.orc_interp.argtypes = [C.c_double, P, P, C.c_int]
            L.orc_interp.restype = C.c_double
            L.orc_refract.argtypes = [P, P, C.c_double, P]
            L.orc_refract.restype = C.c_int
            L.orc_brdf_reflect.argtypes = [P, P, C.c_double, C.c_double]
            L.orc_brdf_reflect.restype = C.c_float
            L.orc_btdf_transmit.argtypes = [P, P, C.c_double, C.c_double, C.c_double, C.c_double]
            L.orc_btdf_transmit.restype = C.c_double
            L.orc_scatter_mult.argtypes = [P, P, P, P, C.c_int]
            L.orc_scatter_mult.restype = C.c_double
            L.orc_emission_dir.argtypes = [P, P, C.c_int32, C.c_uint64, P]
            L.orc_layer_lookup.argtypes = [P, C.c_double, P]
            L.orc_lambert.argtypes = [C.c_double, C.c_double, P]
            L.orc_scatter_direction.argtypes = [P, P, C.c_double, C.c_double, P]
            L.orc_layer_layout.argtypes = [P]
            for name in ("orc_ggx_d_x", "orc_g1_x", "orc_schlick_x"):
                getattr(L, name).argtypes = [C.c_double, C.c_double]
                getattr(L, name).restype = C.c_double
            _lib = L
    return _lib


def _ptr(a: np.ndarray) -> int:
    return a.ctypes.data


# ----------------------------------------------------------------- small wrappers

def philox4x32_10(ctr, key):
    c = np.asarray(ctr, dtype=np.uint32)
    k = np.asarray(key, dtype=np.uint32)
    out = np.zeros(4, dtype=np.uint32)
    lib().orc_philox4x32_10(_ptr(c), _ptr(k), _ptr(out))
    return out


def uniforms(key, photon: int, n: int, stream: int = 0) -> np.ndarray:
    r = abi.make_rng(key, stream)
    return np.array([lib().orc_uniform_at(C.byref(r), photon, j) for j in range(n)])


def rotation_from_z(v) -> np.ndarray:
    vv = np.ascontiguousarray(v, dtype=np.float64)
    R = np.zeros(9, dtype=np.float32)
    lib().orc_rotation_from_z(_ptr(vv), _ptr(R))
    return R.reshape(3, 3)


def interp(x: float, xp: np.ndarray, fp: np.ndarray) -> float:
    xp = np.ascontiguousarray(xp, dtype=np.float64)
    fp = np.ascontiguousarray(fp, dtype=np.float64)
    return lib().orc_interp(float(x), _ptr(xp), _ptr(fp), len(xp))


def refract(d, n, eta):
    dd = np.ascontiguousarray(d, dtype=np.float64)
    nn = np.ascontiguousarray(n, dtype=np.float64)
    out = np.zeros(3)
    tir = lib().orc_refract(_ptr(dd), _ptr(nn), float(eta), _ptr(out))
    return out, bool(tir)


def brdf_reflect(dp, ds, alpha, F0) -> float:
    a = np.ascontiguousarray(dp, dtype=np.float32)
    b = np.ascontiguousarray(ds, dtype=np.float64)
    return float(lib().orc_brdf_reflect(_ptr(a), _ptr(b), float(alpha), float(F0)))


def btdf_transmit(dp, ds, alpha, n_w, F0, albedo) -> float:
    a = np.ascontiguousarray(dp, dtype=np.float32)
    b = np.ascontiguousarray(ds, dtype=np.float64)
    return float(lib().orc_btdf_transmit(_ptr(a), _ptr(b), float(alpha), float(n_w), float(F0), float(albedo)))


def scatter_mult(dp, ds, theta, cdf) -> float:
    a = np.ascontiguousarray(dp, dtype=np.float32)
    b = np.ascontiguousarray(ds, dtype=np.float64)
    return float(lib().orc_scatter_mult(_ptr(a), _ptr(b), _ptr(theta), _ptr(cdf), len(theta)))


def ggx_d(ndh, alpha):
    return lib().orc_ggx_d_x(float(ndh), float(alpha))


def smith_g1(x, alpha):
    return lib().orc_g1_x(float(x), float(alpha))


def schlick(c, F0):
    return lib().orc_schlick_x(float(c), float(F0))


def lambert(u1, u2):
    out = np.zeros(3)
    lib().orc_lambert(float(u1), float(u2), _ptr(out))
    return out


def emission_dirs(model, key, pass_: int, n: int) -> np.ndarray:
    """Starting directions of photons 0..n-1 of one pass (Philox streams)."""
    hs = HostScene(model)
    r = abi.make_rng(key, pass_)
    out = np.zeros((n, 3))
    L = lib()
    for i in range(n):
        L.orc_emission_dir(C.byref(hs.s), C.byref(r), pass_, i, out[i].ctypes.data)
    return out


def scatter_direction(model, din, u1, u2):
    hs = HostScene(model)
    d = np.ascontiguousarray(din, dtype=np.float64)
    out = np.zeros(3)
    lib().orc_scatter_direction(C.byref(hs.s), _ptr(d), float(u1), float(u2), _ptr(out))
    return out


# ----------------------------------------------------------------- scene packing

class HostScene:
    """alb_scene_t with host phase tables, kept alive with the struct."""

    def __init__(self, model, phase_score: np.ndarray | None = None, heat: dict | None = None):
        """heat: optional {"bins", "water_extent", "seafloor_extent"}: host heat
        maps ("radius" uint32, "water" and "floor" FP64, [bins * bins]) in
        self.heat, centred as the HIP path centres them."""
        self.model = model
        self.theta = np.ascontiguousarray(model.phase_theta, dtype=np.float64)
        self.cdf = np.ascontiguousarray(model.phase_cdf, dtype=np.float64)
        self.score = None if phase_score is None else np.ascontiguousarray(phase_score, dtype=np.float64)
        self.s = abi.make_scene(model, _ptr(self.theta), _ptr(self.cdf), 0,
                                _ptr(self.score) if self.score is not None else 0)
        self.layer_arrays = self.layer_structs = None
        if model.layers:
            from pyalbsim_amd.model import layer_arrays

            self.layer_arrays = layer_arrays(model)
            ptrs = [{nm: _ptr(a[nm]) for nm in abi.LAYER_TABLES} for a in self.layer_arrays]
            self.layer_structs = abi.make_layers(model, ptrs)
            self.s.n_layers = len(model.layers)
            self.s.layers = C.addressof(self.layer_structs)
        self.heat = None
        if heat is not None:
            from pyalbsim_amd.footprint import seafloor_intersection, water_surface_intersection

            b = int(heat["bins"])
            self.heat = {"bins": b, "radius": np.zeros(b * b, dtype=np.uint32), "water": np.zeros(b * b),
                         "floor": np.zeros(b * b)}
            self.s.heat_bins = b
            self.s.heat_water_extent = float(heat["water_extent"])
            self.s.heat_floor_extent = float(heat["seafloor_extent"])
            wc, fc = water_surface_intersection(model), seafloor_intersection(model)
            for i in range(3):
                self.s.heat_water_center[i] = float(wc[i])
                self.s.heat_floor_center[i] = float(fc[i])
            self.s.heat_radius = _ptr(self.heat["radius"])
            self.s.heat_water = _ptr(self.heat["water"])
            self.s.heat_floor = _ptr(self.heat["floor"])


def layer_layout() -> tuple:
    """(sizeof, offset of a_tab, offset of tab_n, offset of phase_n) of
    alb_layer_t as the C compiler lays it out."""
    out = np.zeros(4, dtype=np.int64)
    lib().orc_layer_layout(_ptr(out))
    return tuple(int(v) for v in out)


def layer_lookup(model, depths) -> np.ndarray:
    """(layer, velocity, float32 attenuation, float32 albedo) at each depth of a
    layered scene, from the oracle's per-depth lookups."""
    hs = HostScene(model)
    out = np.zeros((len(depths), 4))
    buf = np.zeros(4)
    for i, d in enumerate(np.asarray(depths, dtype=np.float64)):
        lib().orc_layer_lookup(C.byref(hs.s), float(d), _ptr(buf))
        out[i] = buf
    return out


_FIELDS_F32 = ("x", "y", "z", "dx", "dy", "dz", "energy", "time")


def _alloc_sink(capacity: int, query: bool, points: bool = False):
    dt = np.float64 if query else np.float32
    arrs = {}
    for f in _FIELDS_F32:
        if query and f in ("energy", "time"):
            arrs[f] = np.zeros(capacity, dtype=np.float32)
        else:
            arrs[f] = np.zeros(capacity, dtype=dt if f not in ("energy", "time") else np.float32)
    arrs["tag"] = np.zeros(capacity, dtype=np.int64)
    arrs["count"] = np.zeros(1, dtype=np.uint64)
    st = abi.Queries() if query else abi.Records()
    for f in _FIELDS_F32:
        setattr(st, f, _ptr(arrs[f]))
    st.tag = _ptr(arrs["tag"])
    st.capacity = capacity
    st.count = _ptr(arrs["count"])
    if points and not query:
        for f in abi.INTERACTION_FIELDS:
            arrs[f] = np.zeros(capacity, dtype=np.float32)
            setattr(st, f, _ptr(arrs[f]))
    return st, arrs


class _StatsSink(C.Structure):
    _fields_ = [("lam", C.c_void_p), ("gs", C.c_void_p), ("nsc", C.c_void_p), ("nfl", C.c_void_p),
                ("dsc", C.c_void_p), ("rdl", C.c_void_p)]


def trace(model, key, pass_: int, photon_begin: int, n: int, replay: np.ndarray | None = None,
          stream: int | None = None, per_photon: int = 12, stats: bool = False,
          phase_score: np.ndarray | None = None, points: bool = False, heat: dict | None = None,
          stat_layer: int = -1):
    """Trace n photons on the CPU; returns (sinks, counters, n_exhausted).

    sinks[t] is a dict of NumPy arrays (records for the forward pass, queries
    for the backward pass) truncated to the emitted count.  stats: also return
    each record's adjoint path statistics ("lam" water length whose optical
    depth was accumulated, "gs" phase-sampling score sum with the per-segment
    scores phase_score, "nsc" scatters whose omega0 was applied, "dsc"
    completed free paths ended by a scatter, "nfl" floor reflections before
    it, and "rdl" 1 for a scatter record whose own free path counts).
    stat_layer >= 0 counts the water length and the scatters of that layer of
    a layered column only (a single constant layer is layer 0).
    points: forward records also carry the interaction points (fwi_*, sfi_*).
    heat: optional {"bins", "water_extent", "seafloor_extent"}; receives the
    forward scatter-radius counts under "radius" ([bins, bins]).
    """
    hs = HostScene(model, phase_score, heat)
    rep = None
    if replay is not None:
        rep = np.ascontiguousarray(replay, dtype=np.float64)
        assert rep.shape[0] == n
    stream = pass_ if stream is None else stream
    rng = abi.make_rng(key, stream, _ptr(rep) if rep is not None else 0,
                       rep.shape[1] if rep is not None else 0)
    cap = max(16, per_photon * n)
    while True:
        query = pass_ == abi.PASS_BACKWARD
        sinks = [_alloc_sink(cap, query, points) for _ in range(abi.N_TYPES)]
        if hs.heat is not None:
            hs.heat["radius"][:] = 0
        rec_arr = (abi.Records * 4)()
        q_arr = (abi.Queries * 4)()
        for t in range(4):
            if query:
                q_arr[t] = sinks[t][0]
                rec_arr[t] = abi.Records()
            else:
                rec_arr[t] = sinks[t][0]
                q_arr[t] = abi.Queries()
        ctr = np.zeros(abi.N_COUNTERS, dtype=np.uint64)
        if stats:
            st_arrs = [{"lam": np.zeros(cap), "gs": np.zeros(cap), "nsc": np.zeros(cap, dtype=np.int32),
                        "nfl": np.zeros(cap, dtype=np.int32), "dsc": np.zeros(cap, dtype=np.int32),
                        "rdl": np.zeros(cap, dtype=np.int32)} for _ in range(4)]
            ss = (_StatsSink * 4)()
            for t in range(4):
                for f in ("lam", "gs", "nsc", "nfl", "dsc", "rdl"):
                    setattr(ss[t], f, _ptr(st_arrs[t][f]))
            ex = lib().orc_trace_stats(C.byref(hs.s), C.byref(rng), pass_, photon_begin, n,
                                       C.cast(rec_arr, C.c_void_p), C.cast(q_arr, C.c_void_p), _ptr(ctr),
                                       C.cast(ss, C.c_void_p), int(stat_layer))
        else:
            ex = lib().orc_trace(C.byref(hs.s), C.byref(rng), pass_, photon_begin, n,
                                 C.cast(rec_arr, C.c_void_p), C.cast(q_arr, C.c_void_p), _ptr(ctr))
        counts = [int(s[1]["count"][0]) for s in sinks]
        if max(counts) <= cap:
            break
        cap = max(counts) + 16
    out = []
    for t in range(4):
        arrs = sinks[t][1]
        c = counts[t]
        d = {f: arrs[f][:c].copy() for f in (*_FIELDS_F32, "tag")}
        if points and not query:
            d.update({f: arrs[f][:c].copy() for f in abi.INTERACTION_FIELDS})
        if stats:
            d.update({f: st_arrs[t][f][:c].copy() for f in ("lam", "gs", "nsc", "nfl", "dsc", "rdl")})
        out.append(d)
    if heat is not None:
        b = hs.heat["bins"]
        heat["radius"] = hs.heat["radius"].reshape(b, b).copy()
    return out, ctr, int(ex)


def knn(x, y, z, queries: np.ndarray, k: int):
    """Brute-force k-NN ordered by (squared distance, index): (idx, d2)."""
    x = np.ascontiguousarray(x, dtype=np.float32)
    y = np.ascontiguousarray(y, dtype=np.float32)
    z = np.ascontiguousarray(z, dtype=np.float32)
    q = np.ascontiguousarray(queries, dtype=np.float64)
    nq = q.shape[0]
    idx = np.zeros((nq, k), dtype=np.int32)
    d2 = np.zeros((nq, k), dtype=np.float64)
    L = lib()
    for j in range(nq):
        rc = L.orc_knn_one(_ptr(x), _ptr(y), _ptr(z), len(x), q[j, 0], q[j, 1], q[j, 2], k,
                           idx[j].ctypes.data, d2[j].ctypes.data)
        if rc != 0:
            raise IndexError("k exceeds the photon map size")
    return idx, d2


def gather(model, ptype: int, mp: dict, qs: dict, n_bins: int | None = None,
           want_nbrs: bool = False, heat: dict | None = None, nbrs: np.ndarray | None = None):
    """FP64 waveform of one photon type from map arrays mp and queries qs.
    heat: optional {"bins", "water_extent", "seafloor_extent"} (mp must hold
    the interaction points); receives this gather's "water" and "floor" heat
    maps ([bins, bins], FP64).  nbrs: optional [nq, k] neighbour indices from
    an exact k-NN (e.g. a k-d tree), used instead of the brute-force search."""
    hs = HostScene(model, None, heat)
    pts = [None] * 4
    if heat is not None:
        pts = [np.ascontiguousarray(mp[f], dtype=np.float32) for f in ("fwi_x", "fwi_z", "sfi_x", "sfi_z")]
    nb_in = None
    if nbrs is not None:
        nb_in = np.ascontiguousarray(nbrs, dtype=np.int32)
        assert nb_in.shape == (len(qs["x"]), model.k)
    n_bins = model.n_bins if n_bins is None else n_bins
    m = {f: np.ascontiguousarray(mp[f], dtype=np.float32) for f in _FIELDS_F32}
    nq = len(qs["x"])
    qa = {f: np.ascontiguousarray(qs[f], dtype=np.float64 if f not in ("energy", "time") else np.float32)
          for f in _FIELDS_F32}
    qst = abi.Queries()
    for f in _FIELDS_F32:
        setattr(qst, f, _ptr(qa[f]))
    qst.capacity = nq
    wf = np.zeros(n_bins, dtype=np.float64)
    nbr = np.zeros((nq, model.k), dtype=np.int32) if want_nbrs else None
    rad = np.zeros(nq, dtype=np.float64) if want_nbrs else None
    rc = lib().orc_gather_ex(C.byref(hs.s), ptype, *[_ptr(m[f]) for f in _FIELDS_F32], len(m["x"]),
                             C.byref(qst), nq, _ptr(wf), n_bins,
                             _ptr(nbr) if want_nbrs else None, _ptr(rad) if want_nbrs else None,
                             _ptr(nb_in) if nb_in is not None else None,
                             *[_ptr(a) if a is not None else None for a in pts])
    if rc == -2:
        raise IndexError("a given neighbour index lies outside the map")
    if rc != 0:
        if len(m["x"]) == 0:
            raise KeyError(ptype)
        raise IndexError("k exceeds the photon map size")
    if heat is not None:
        b = hs.heat["bins"]
        heat["water"] = hs.heat["water"].reshape(b, b).copy()
        heat["floor"] = hs.heat["floor"].reshape(b, b).copy()
    if want_nbrs:
        return wf, nbr, rad
    return wf


def radius_gather(model, ptype: int, mp: dict, qs: dict, radius: float, n_bins: int | None = None):
    """Fixed-radius gather, brute force: (FP64 waveform, per-query record
    counts, per-query sums of record indices)."""
    hs = HostScene(model)
    n_bins = model.n_bins if n_bins is None else n_bins
    m = {f: np.ascontiguousarray(mp[f], dtype=np.float32) for f in _FIELDS_F32}
    nq = len(qs["x"])
    qa = {f: np.ascontiguousarray(qs[f], dtype=np.float64 if f not in ("energy", "time") else np.float32)
          for f in _FIELDS_F32}
    qst = abi.Queries()
    for f in _FIELDS_F32:
        setattr(qst, f, _ptr(qa[f]))
    qst.capacity = nq
    wf = np.zeros(n_bins, dtype=np.float64)
    cnt = np.zeros(nq, dtype=np.int64)
    isum = np.zeros(nq, dtype=np.int64)
    lib().orc_radius_gather(C.byref(hs.s), ptype, *[_ptr(m[f]) for f in _FIELDS_F32], len(m["x"]),
                            C.byref(qst), nq, float(radius), _ptr(wf), n_bins, _ptr(cnt), _ptr(isum))
    return wf, cnt, isum


def run(model, n_forward: int, n_backward: int, key, replay_fwd=None, replay_bwd=None):
    """Whole path on the CPU: forward maps, backward queries, FP64 waveforms."""
    maps, cf, exf = trace(model, key, abi.PASS_FORWARD, 0, n_forward, replay_fwd)
    qs, cb, exb = trace(model, key, abi.PASS_BACKWARD, 0, n_backward, replay_bwd)
    wfs = []
    for t in range(4):
        if len(qs[t]["x"]) == 0:
            wfs.append(np.zeros(model.n_bins))
            continue
        if len(maps[t]["x"]) == 0:
            raise KeyError(t)
        wfs.append(gather(model, t, maps[t], qs[t]))
    return {"waveforms": wfs, "maps": maps, "queries": qs, "counters_fwd": cf, "counters_bwd": cb,
            "exhausted": exf + exb}
