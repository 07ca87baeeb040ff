"""Host runtime over the HIP library: device buffers (torch tensors used only
as allocations), ctypes calls into libalbsim_hip.so, capacity handling.

There is no CPU fallback: every call requires the built HIP library and a
GPU, and raises otherwise.
"""
from __future__ import annotations

import ctypes as C
import os
from dataclasses import dataclass, field

import numpy as np

from . import abi
from .model import SceneModel

LIB_PATH = os.path.join(os.path.dirname(os.path.abspath(__file__)), "_lib", "libalbsim_hip.so")
_lib = None

# header entry points, checked by the ABI test
EXPORTS = ["alb_abi_version", "alb_scene_size", "alb_last_error", "alb_trace", "alb_map_workspace_size",
           "alb_map_build", "alb_gather_workspace_size", "alb_gather", "alb_gather_order", "alb_knn",
           "alb_contrib", "alb_trace_batched", "alb_map_build_batched", "alb_gather_order_batched",
           "alb_knn_batched", "alb_contrib_batched", "alb_adjoint_contrib", "alb_replay_index_workspace_size",
           "alb_replay_index_build", "alb_replay_batched", "alb_radius_gather", "alb_gather_static"]


class AlbError(RuntimeError):
    pass


def check_k(models) -> None:
    """Raise AlbError before any device work when a scene's k exceeds what the k-NN kernels take."""
    for m in models:
        if not 1 <= int(m.k) <= abi.MAX_K:
            raise AlbError(f"photon_mapping_k = {m.k} outside [1, {abi.MAX_K}], the k-NN gather's range")


def load_library(path: str = LIB_PATH):
    """Load libalbsim_hip.so and declare its C signatures (no GPU needed)."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise AlbError(f"HIP library not built: {path} is missing (run python -m pyalbsim_amd.build)")
    L = C.CDLL(path)
    P = C.c_void_p
    L.alb_abi_version.restype = C.c_int
    L.alb_scene_size.restype = C.c_size_t
    L.alb_last_error.restype = C.c_char_p
    L.alb_trace.argtypes = [P, P, C.c_int32, C.c_int64, C.c_int64, P, P, P, P]
    L.alb_trace.restype = C.c_int
    L.alb_map_workspace_size.argtypes = [C.c_int64, C.POINTER(C.c_size_t)]
    L.alb_map_workspace_size.restype = C.c_int
    L.alb_map_build.argtypes = [P, C.c_int64, C.c_int32, P, P, C.c_size_t, P]
    L.alb_map_build.restype = C.c_int
    L.alb_gather_workspace_size.argtypes = [C.c_int64, C.c_int32, C.POINTER(C.c_size_t)]
    L.alb_gather_workspace_size.restype = C.c_int
    L.alb_gather.argtypes = [P, P, C.c_int32, P, C.c_int64, C.c_int64, P, C.c_int64, C.c_int32, C.c_int32,
                             P, P, P, C.c_size_t, P, P]
    L.alb_gather.restype = C.c_int
    L.alb_gather_static.argtypes = [P, P, C.c_int32, P, C.c_int64, C.c_int64, P, C.c_int64, C.c_int32, C.c_int32,
                                    P, P, P, P]
    L.alb_gather_static.restype = C.c_int
    L.alb_gather_order.argtypes = [P, P, C.c_int64, C.c_int64, P, P, C.c_size_t, P]
    L.alb_radius_gather.argtypes = [P, P, C.c_int32, P, C.c_int64, C.c_int64, P, C.c_double, P, C.c_int64,
                                    C.c_int32, C.c_int32, P, P, P, P]
    L.alb_radius_gather.restype = C.c_int
    L.alb_gather_order.restype = C.c_int
    L.alb_knn.argtypes = [P, C.c_int32, C.c_int32, P, C.c_int64, P, C.c_int64, C.c_int64, C.c_int32, P, P, P, P, P, P]
    L.alb_knn.restype = C.c_int
    L.alb_contrib.argtypes = [P, P, C.c_int32, P, C.c_int64, P, C.c_int64, C.c_int64, C.c_int32, P, P, P, C.c_int64,
                              C.c_int32, C.c_int32, P, C.c_int32, P]
    L.alb_contrib.restype = C.c_int
    L.alb_trace_batched.argtypes = [P, C.c_int32, P, P, C.c_int32, P, P, P, C.c_int32, C.c_int32, P, P]
    L.alb_trace_batched.restype = C.c_int
    L.alb_map_build_batched.argtypes = [P, C.c_int64, C.c_int32, P, P, P, C.c_size_t, P]
    L.alb_map_build_batched.restype = C.c_int
    L.alb_gather_order_batched.argtypes = [P, P, P, C.c_int64, P, P, C.c_size_t, P]
    L.alb_gather_order_batched.restype = C.c_int
    L.alb_knn_batched.argtypes = [P, P, C.c_int32, C.c_int32, P, P, C.c_int64, C.c_int64, P, P, P, P, P, P]
    L.alb_knn_batched.restype = C.c_int
    L.alb_contrib_batched.argtypes = [P, C.c_int32, C.c_int32, P, C.c_int32, P, P, C.c_int64, C.c_int64, P, P, P,
                                      P, P, C.c_int32, P, C.c_int32, C.c_int64, P]
    L.alb_contrib_batched.restype = C.c_int
    L.alb_adjoint_contrib.argtypes = [P, C.c_int32, C.c_int32, P, C.c_int32, P, P, C.c_int64, C.c_int64, P, P, P, P,
                                      P, P, P, P]
    L.alb_adjoint_contrib.restype = C.c_int
    L.alb_replay_index_workspace_size.argtypes = [C.c_int64, C.POINTER(C.c_size_t)]
    L.alb_replay_index_workspace_size.restype = C.c_int
    L.alb_replay_index_build.argtypes = [P, P, C.c_int64, P, P, P, P, C.c_size_t, P]
    L.alb_replay_index_build.restype = C.c_int
    L.alb_replay_batched.argtypes = [P, C.c_int32, P, P, C.c_int32, P, P, C.c_int32, C.c_int32, P, P, C.c_int32, P]
    L.alb_replay_batched.restype = C.c_int
    if L.alb_abi_version() != abi.ALB_ABI_VERSION:
        raise AlbError("libalbsim_hip.so ABI version mismatch")
    if L.alb_scene_size() != C.sizeof(abi.Scene):
        raise AlbError("alb_scene_t layout mismatch between include/alb_abi.h and pyalbsim_amd/abi.py")
    _lib = L
    return L


def _check(rc: int, what: str):
    if rc != abi.ALB_OK:
        msg = load_library().alb_last_error().decode(errors="replace")
        if rc == abi.ALB_ERR_CAPACITY:
            raise MemoryError(f"{what}: {msg}")
        raise AlbError(f"{what} failed ({rc}): {msg}")


def _torch():
    import torch

    if not torch.cuda.is_available():
        raise AlbError("no GPU available: the photon-tracing path runs only on the HIP device")
    return torch


def _stream_ptr():
    torch = _torch()
    return torch.cuda.current_stream().cuda_stream


def _p(t) -> int:
    return int(t.data_ptr()) if t is not None else 0


def _upload_layers(model: SceneModel, device):
    """Layer tables and the alb_layer_t array of a layered scene on the device:
    (device struct tensor or None, tensors to keep alive)."""
    from .model import layer_arrays

    if not model.layers:
        return None, []
    torch = _torch()
    keep, ptrs = [], []
    for arrs in layer_arrays(model):
        tens = {nm: torch.as_tensor(arrs[nm], device=device) for nm in abi.LAYER_TABLES}
        keep.extend(tens.values())
        ptrs.append({nm: _p(t) for nm, t in tens.items()})
    structs = abi.make_layers(model, ptrs)
    dev = torch.as_tensor(np.frombuffer(bytes(structs), dtype=np.uint8).copy(), device=device)
    keep.append(dev)
    return dev, keep


class DeviceScene:
    """Scene constants with the phase tables resident on the device."""

    def __init__(self, model: SceneModel, device="cuda"):
        torch = _torch()
        self.model = model
        self.theta = torch.as_tensor(np.ascontiguousarray(model.phase_theta), dtype=torch.float64, device=device)
        self.cdf = torch.as_tensor(np.ascontiguousarray(model.phase_cdf), dtype=torch.float64, device=device)
        from .model import phase_bin_mass

        self.mass = torch.as_tensor(phase_bin_mass(model.phase_theta, model.phase_cdf), dtype=torch.float64,
                                    device=device)
        self.s = abi.make_scene(model, _p(self.theta), _p(self.cdf), _p(self.mass))
        self.heat = None
        self.layers, self._layer_keep = _upload_layers(model, device)
        if self.layers is not None:
            self.s.n_layers = len(model.layers)
            self.s.layers = _p(self.layers)
        self.device = device

    def enable_heatmaps(self, bins: int, water_extent: float, seafloor_extent: float):
        """Allocate the sampled heat maps (core/simulation.py:51-69): water and
        seafloor energy [bins, bins] (FP64) and the forward scatter-radius
        counts [bins, bins], centred on the laser's water-surface and seafloor
        intersections (physics/models/simulation.py:90-124)."""
        from .footprint import seafloor_intersection, water_surface_intersection

        torch = _torch()
        b = int(bins)
        if b <= 0 or not (water_extent > 0 and seafloor_extent > 0):
            raise ValueError("heat maps need bins > 0 and positive extents")
        self.heat = {"bins": b,
                     "radius": torch.zeros(b * b, dtype=torch.int32, device=self.device),
                     "water": torch.zeros(b * b, dtype=torch.float64, device=self.device),
                     "floor": torch.zeros(b * b, dtype=torch.float64, device=self.device)}
        s = self.s
        s.heat_bins = b
        s.heat_water_extent = float(water_extent)
        s.heat_floor_extent = float(seafloor_extent)
        wc, fc = water_surface_intersection(self.model), seafloor_intersection(self.model)
        for i in range(3):
            s.heat_water_center[i] = float(wc[i])
            s.heat_floor_center[i] = float(fc[i])
        s.heat_radius = _p(self.heat["radius"])
        s.heat_water = _p(self.heat["water"])
        s.heat_floor = _p(self.heat["floor"])

    def clear_heatmaps(self):
        if self.heat:
            for f in ("radius", "water", "floor"):
                self.heat[f].zero_()

    def heatmaps(self) -> dict:
        """Host copies: water and seafloor energy, scatter-radius counts, each [bins, bins]."""
        b = self.heat["bins"]
        return {f: self.heat[f].cpu().numpy().reshape(b, b) for f in ("water", "floor", "radius")}


REC_F32 = ("x", "y", "z", "dx", "dy", "dz", "energy", "time")


@dataclass
class Sink:
    """Device SoA sink for records (float32 geometry) or queries (float64 geometry)."""

    tensors: dict
    capacity: int
    query: bool
    struct: object = None

    @staticmethod
    def alloc(capacity: int, query: bool, tags: bool, device="cuda", scenes: bool = False,
              points: bool = False) -> "Sink":
        torch = _torch()
        capacity = max(int(capacity), 1)
        gdt = torch.float64 if query else torch.float32
        t = {}
        for f in REC_F32:
            t[f] = torch.empty(capacity, dtype=torch.float32 if f in ("energy", "time") else gdt, device=device)
        t["tag"] = torch.empty(capacity, dtype=torch.int64, device=device) if tags else None
        t["scene"] = torch.empty(capacity, dtype=torch.int32, device=device) if scenes else None
        for f in abi.INTERACTION_FIELDS:  # first water / seafloor interaction points (heat maps)
            t[f] = torch.empty(capacity, dtype=torch.float32, device=device) if points and not query else None
        t["count"] = torch.zeros(1, dtype=torch.int64, device=device)
        st = abi.Queries() if query else abi.Records()
        for f in REC_F32:
            setattr(st, f, _p(t[f]))
        st.tag = _p(t["tag"])
        st.scene = _p(t["scene"])
        for f in abi.INTERACTION_FIELDS:
            setattr(st, f, _p(t[f]))
        st.capacity = capacity
        st.count = _p(t["count"])
        return Sink(t, capacity, query, st)

    def view(self, n: int) -> dict:
        return {f: (self.tensors[f][:n] if self.tensors.get(f) is not None else None)
                for f in (*REC_F32, "tag", "scene", *abi.INTERACTION_FIELDS)}


# expected records per photon and type (BOTTOM, SCATTER, SURF_REFL, SURF_TRANS) before a pilot
_DEFAULT_RATES = (1.5, 8.0, 1.0, 1.5)


def trace(dscene: DeviceScene, key, pass_: int, n: int, begin: int = 0, replay=None, tags: bool = False,
          rates=None, counters=None, stream=None, pilot: int = 65536, scene_id: int = 0, points: bool = False,
          _count_heat: bool = True):
    """Trace photons [begin, begin + n) of one pass on the GPU.

    Returns (sinks, counts): sinks[t] is a Sink holding counts[t] items.
    Capacities come from a pilot run over the first photons (same counters, so
    it does not change the result); an overflow reruns the whole range with
    exact capacities.  replay: device float64 tensor [n, L] of uniforms.
    points: forward records also carry the photons' first water and seafloor
    interaction points (heat maps, PhotonMapData).  A scene with heat maps
    counts the forward seafloor hits of this range into its radius histogram
    once (not for the pilot, and not twice on a rerun).
    """
    torch = _torch()
    L = load_library()
    query = pass_ == abi.PASS_BACKWARD
    rp = _p(replay) if replay is not None else 0
    rlen = int(replay.shape[1]) if replay is not None else 0
    rng = abi.make_rng(key, pass_ | (int(scene_id) << 1), rp, rlen)
    s = stream if stream is not None else _stream_ptr()
    if rates is None and n > 4 * pilot and replay is None:
        sinks, counts = trace(dscene, key, pass_, pilot, begin, None, False, _DEFAULT_RATES, None, stream, pilot,
                              scene_id, _count_heat=False)
        rates = [max(c / pilot, 0.05) for c in counts]
        del sinks
    rates = rates or _DEFAULT_RATES
    caps = [int(r * n * 1.05) + 1024 for r in rates]
    scene = dscene.s
    if dscene.heat is not None and not (_count_heat and pass_ == abi.PASS_FORWARD):
        scene = abi.Scene.from_buffer_copy(dscene.s)
        scene.heat_radius = None
    radius0 = dscene.heat["radius"].clone() if (dscene.heat is not None and scene is dscene.s) else None
    for attempt in range(2):
        sinks = [Sink.alloc(caps[t], query, tags, points=points) for t in range(abi.N_TYPES)]
        arr_r = (abi.Records * 4)()
        arr_q = (abi.Queries * 4)()
        for t in range(4):
            if query:
                arr_q[t] = sinks[t].struct
            else:
                arr_r[t] = sinks[t].struct
        ctr_t = counters if counters is not None else None
        if ctr_t is not None and attempt == 1:
            ctr_t.zero_()
        if radius0 is not None and attempt == 1:
            dscene.heat["radius"].copy_(radius0)
        rc = L.alb_trace(C.byref(scene), C.byref(rng), pass_, begin, n,
                         C.cast(arr_r, C.c_void_p) if not query else None,
                         C.cast(arr_q, C.c_void_p) if query else None, _p(ctr_t), C.c_void_p(s))
        _check(rc, "alb_trace")
        counts = [int(x) for x in torch.cat([sk.tensors["count"] for sk in sinks]).cpu().tolist()]
        if all(counts[t] <= caps[t] for t in range(4)):
            return sinks, counts
        caps = [max(counts[t], 1) for t in range(4)]
    raise AlbError("alb_trace: capacity overflow persisted after resizing")


@dataclass
class DeviceMap:
    n: int
    dims: int
    tensors: dict = field(default_factory=dict)
    struct: object = None
    index: object = None  # abi.SceneIndex of a map of many scenes


def build_map(records: Sink, n: int, dims: int, stream=None) -> DeviceMap:
    """Morton-sort n records of a forward sink into a map with a radix tree."""
    torch = _torch()
    L = load_library()
    dev = records.tensors["x"].device
    m = DeviceMap(n=n, dims=dims)
    t = m.tensors
    for f in REC_F32:
        t[f] = torch.empty(max(n, 1), dtype=torch.float32, device=dev)
    t["perm"] = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    t["keys"] = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    t["meta"] = torch.zeros(8, dtype=torch.float64, device=dev)
    ni = max(n - 1, 1)
    t["nodes"] = torch.empty((ni, 16), dtype=torch.int32, device=dev)  # 64-byte nodes
    st = abi.Map()
    st.n = n
    st.dims = dims
    for f in ("x", "y", "z", "dx", "dy", "dz", "energy", "time", "perm", "keys", "meta", "nodes"):
        setattr(st, f, _p(t[f]))
    if records.tensors.get("fwi_x") is not None:  # interaction points follow the sort
        for f in abi.INTERACTION_FIELDS:
            t[f] = torch.empty(max(n, 1), dtype=torch.float32, device=dev)
            setattr(st, f, _p(t[f]))
    m.struct = st
    if n == 0:
        return m
    ws_bytes = C.c_size_t(0)
    _check(L.alb_map_workspace_size(n, C.byref(ws_bytes)), "alb_map_workspace_size")
    ws = torch.empty(ws_bytes.value, dtype=torch.uint8, device=dev)
    s = stream if stream is not None else _stream_ptr()
    _check(L.alb_map_build(C.byref(records.struct), n, dims, C.byref(st), _p(ws), ws_bytes.value, C.c_void_p(s)),
           "alb_map_build")
    m.tensors["_ws"] = ws  # keep alive until the stream has consumed it
    return m


CHUNK = 1 << 24  # queries per k-NN / contribution round at k = 100 (2^22: SCATTER k-NN 3 % slower; 2^25: the same step time; rows take 4 k B per query)


def gather_chunk(q_n: int, k: int) -> int:
    """Queries per k-NN / contribution round: CHUNK at k = 100, fewer for
    larger k so that the neighbour rows stay near 6.7 GB."""
    return min(q_n, CHUNK if k <= 100 else max(1, CHUNK * 100 // k))
# LDS-privatised bins of the contribution kernels.  Most (query, neighbour)
# pairs fall in the first ~4000 bins after the surface return (backward paths
# scatter on long after it), so a window of this many bins from the gate's
# start takes ~99 % of their FP64 adds into LDS; a 1000-bin window took
# 5-12 % and left the rest to global atomics.  The window changes the order
# of the FP64 additions only.  The batched sweep's deeper, more turbid scenes
# have later tails: their window is 16384 bins (128 KB of LDS, one 1024-thread
# block per CU; contribution 251 -> 177 ms at 7680 bins, 108 ms at 16384 on
# 256 sweep scenes); the single shot gains nothing beyond 4096.
LDS_WINDOW = 4096
LDS_WINDOW_BATCHED = 16384


def lds_window(n_bins: int, win0: int, win_len: int, width: int = None) -> tuple:
    """The gate (win0, win_len) widened to width (default LDS_WINDOW) bins
    within the waveform."""
    w = LDS_WINDOW if width is None else width
    return int(win0), int(max(win_len, min(w, n_bins - win0)))
REPLICAS = 64    # waveform copies that spread the global FP64 atomics of one gather
BATCHED_REPLICAS = 0   # the same for a batched gather (copies of one type's waveforms of every scene)


def gather(dscene: DeviceScene, dmap: DeviceMap, ptype: int, queries: Sink, q_n: int, waveform, window,
           q_begin: int = 0, nbr_out=None, radius_out=None, counters=None, stream=None, events=None):
    """Accumulate the k-NN gather of q_n queries of one type into waveform (float64 device tensor).

    Runs alb_gather_order, then alb_knn + alb_contrib per chunk of CHUNK
    queries.  nbr_out / radius_out (optional, given together) receive the
    neighbour sets in query order, (distance, index) sorted.  events: optional
    dict collecting (start, stop) HIP event pairs of the knn and contrib
    launches under keys "knn" and "contrib".
    """
    torch = _torch()
    L = load_library()
    if q_n == 0:
        return
    k = dscene.model.k
    win0, win_len = lds_window(dscene.model.n_bins, *window)
    s = stream if stream is not None else _stream_ptr()
    dev = waveform.device
    direct = nbr_out is not None
    if direct != (radius_out is not None):
        raise ValueError("nbr_out and radius_out go together")
    b = C.c_size_t(0)
    _check(L.alb_gather_workspace_size(q_n, 1, C.byref(b)), "alb_gather_workspace_size")
    ws = torch.empty(b.value, dtype=torch.uint8, device=dev)
    order = torch.empty(q_n, dtype=torch.int32, device=dev)
    _check(L.alb_gather_order(C.byref(dmap.struct), C.byref(queries.struct), q_begin, q_n, _p(order), _p(ws),
                              b.value, C.c_void_p(s)), "alb_gather_order")
    chunk = q_n if direct else gather_chunk(q_n, k)
    knorm = torch.empty(q_n if direct else chunk, dtype=torch.float64, device=dev)
    n_rep = REPLICAS
    rep = torch.empty(n_rep * int(waveform.numel()), dtype=torch.float64, device=dev) if n_rep else None
    nbr = nbr_out if direct else torch.empty((chunk, k), dtype=torch.int32, device=dev)
    rad = radius_out if direct else torch.empty(chunk, dtype=torch.float64, device=dev)

    def ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record(torch.cuda.current_stream())
        return e

    work = torch.empty(abi.KNN_WORK_WORDS, dtype=torch.int64, device=dev)  # cleared by each alb_knn call
    for c0 in range(0, q_n, chunk):
        cn = min(chunk, q_n - c0)
        e0 = ev() if events is not None else None
        _check(L.alb_knn(C.byref(dmap.struct), ptype, k, C.byref(queries.struct), q_begin, _p(order), c0, cn,
                         1 if direct else 0, _p(nbr), _p(rad), _p(knorm), _p(counters), _p(work), C.c_void_p(s)),
               "alb_knn")
        e1 = ev() if events is not None else None
        _check(L.alb_contrib(C.byref(dscene.s), C.byref(dmap.struct), ptype, C.byref(queries.struct), q_begin,
                             _p(order), c0, cn, 1 if direct else 0, _p(nbr), _p(knorm), _p(waveform),
                             int(waveform.numel()), int(win0), int(win_len), _p(rep), n_rep, C.c_void_p(s)),
               "alb_contrib")
        if events is not None:
            e2 = ev()
            events.setdefault("knn", []).append((e0, e1))
            events.setdefault("contrib", []).append((e1, e2))
    return (ws, order, knorm, nbr, rad, rep)  # the caller keeps these alive until the stream has run


MAP_DIMS = (2, 3, 2, 2)  # BOTTOM and the two surface maps are planar; SCATTER is volumetric


def radius_gather(dscene: DeviceScene, dmap: DeviceMap, ptype: int, queries: Sink, q_n: int, radius: float,
                  waveform, window, q_begin: int = 0, sort: bool = True, counts: bool = False, counters=None,
                  stream=None, check: bool = True):
    """Accumulate the fixed-radius gather of q_n queries of one type into
    waveform (float64 device tensor): every map record within radius of a
    query contributes with the kernel 1/(pi r^2) (1/((4/3) pi r^3) for the
    scatter map).  sort: visit the queries in Morton order (alb_gather_order).
    counts: also return per-query record counts and index sums (int64 device
    tensors).  check: synchronise and raise AlbError when a traversal
    overflowed (its sum would be incomplete).  Returns (count, index_sum,
    keep), keep holding the buffers the launch uses."""
    torch = _torch()
    L = load_library()
    win0, win_len = window  # its pairs fall in the gate; a wider window only costs occupancy
    s = stream if stream is not None else _stream_ptr()
    dev = waveform.device
    order = ws = None
    if sort and q_n > 0:
        b = C.c_size_t(0)
        _check(L.alb_gather_workspace_size(q_n, 1, C.byref(b)), "alb_gather_workspace_size")
        ws = torch.empty(b.value, dtype=torch.uint8, device=dev)
        order = torch.empty(q_n, dtype=torch.int32, device=dev)
        _check(L.alb_gather_order(C.byref(dmap.struct), C.byref(queries.struct), q_begin, q_n, _p(order), _p(ws),
                                  b.value, C.c_void_p(s)), "alb_gather_order")
    cnt = torch.zeros(max(q_n, 1), dtype=torch.int64, device=dev) if counts else None
    isum = torch.zeros(max(q_n, 1), dtype=torch.int64, device=dev) if counts else None
    ctr = counters if counters is not None else (torch.zeros(abi.N_COUNTERS, dtype=torch.int64, device=dev)
                                                 if check else None)
    ovf0 = int(ctr[abi.CTR_OVERFLOW].item()) if (check and counters is not None) else 0
    _check(L.alb_radius_gather(C.byref(dscene.s), C.byref(dmap.struct), ptype, C.byref(queries.struct), q_begin,
                               q_n, _p(order), float(radius), _p(waveform), int(waveform.numel()), int(win0),
                               int(win_len), _p(cnt), _p(isum), _p(ctr), C.c_void_p(s)), "alb_radius_gather")
    if check and q_n > 0:
        if int(ctr[abi.CTR_OVERFLOW].item()) != ovf0:
            raise AlbError("alb_radius_gather: a query's traversal stack overflowed")
    return cnt, isum, (order, ws, ctr)


# ----------------------------------------------------------------- many scenes

class DeviceSceneTable:
    """alb_scene_t[n_scenes] on the device; scenes with equal phase tables share them."""

    def __init__(self, models, device="cuda", adjoint: bool = False, gather_phase=None):
        """adjoint: also build the phase-derivative tables of HG scenes.
        gather_phase: optional models whose phase tables the gather's phase
        mass is taken from (default: each scene's own)."""
        torch = _torch()
        self.models = list(models)
        self.n = len(self.models)
        if self.n == 0:
            raise ValueError("no scenes")
        self.adjoint = adjoint
        self.device = device
        self.tables = {}
        self.layer_tables = {}
        raw = self._pack(self.models, gather_phase)
        self.dev = torch.as_tensor(raw, device=device)

    def _pack(self, models, gather_phase=None) -> np.ndarray:
        """alb_scene_t bytes of models; phase and layer tables are uploaded
        once per distinct table and kept by the table object."""
        import hashlib

        from .model import phase_bin_mass

        torch = _torch()
        device = self.device
        ks = {m.k for m in models}
        if len(ks) != 1:
            raise ValueError("all scenes of a batch must use the same photon_mapping_k")
        self.k = ks.pop()
        structs = (abi.Scene * len(models))()
        for i, m in enumerate(models):
            gm = gather_phase[i] if gather_phase is not None else m
            h = hashlib.sha1(np.ascontiguousarray(m.phase_cdf).tobytes()
                             + np.ascontiguousarray(m.phase_theta).tobytes()
                             + np.ascontiguousarray(gm.phase_cdf).tobytes()).hexdigest()
            if h not in self.tables:
                th = torch.as_tensor(np.ascontiguousarray(m.phase_theta), dtype=torch.float64, device=device)
                cd = torch.as_tensor(np.ascontiguousarray(m.phase_cdf), dtype=torch.float64, device=device)
                ms = torch.as_tensor(phase_bin_mass(gm.phase_theta, gm.phase_cdf), dtype=torch.float64, device=device)
                sc = dm = None
                if self.adjoint and m.phase_g is not None:
                    from .adjoint import phase_g_tables

                    score, dlm = phase_g_tables(m.phase_g)
                    sc = torch.as_tensor(score, dtype=torch.float64, device=device)
                    dm = torch.as_tensor(dlm, dtype=torch.float64, device=device)
                self.tables[h] = (th, cd, ms, sc, dm)
            th, cd, ms, sc, dm = self.tables[h]
            structs[i] = abi.make_scene(m, _p(th), _p(cd), _p(ms), _p(sc), _p(dm))
            if m.layers:
                if self.adjoint and any(lm.profiled for lm in m.layers):
                    raise NotImplementedError("the adjoint covers columns of constant layers (no depth profiles)")
                if id(m) not in self.layer_tables:
                    self.layer_tables[id(m)] = (m, _upload_layers(m, device))
                structs[i].n_layers = len(m.layers)
                structs[i].layers = _p(self.layer_tables[id(m)][1][0])
        self.host = structs
        self.n_bins = np.array([m.n_bins for m in models], dtype=np.int64)
        return np.frombuffer(bytes(structs), dtype=np.uint8).copy()

    def update(self, models, gather_phase=None):
        """Rewrite the scene structs in place (same number of scenes), so that
        launches captured in a HIP graph see the new scenes on replay."""
        torch = _torch()
        models = list(models)
        if len(models) != self.n:
            raise ValueError("update needs the same number of scenes")
        raw = self._pack(models, gather_phase)
        self.dev.copy_(torch.as_tensor(raw))
        self.models = models


def _dev_i64(values, device="cuda"):
    torch = _torch()
    return torch.as_tensor(np.asarray(values, dtype=np.int64), device=device)


def trace_batched(table: DeviceSceneTable, key, pass_: int, counts, tags: bool = False, counters=None,
                  block: int = 256, refill: int = 4, stream=None, pilot_fraction: float = 1 / 64,
                  caps=None, stream_base: int = 0):
    """Trace counts[s] photons of every scene s in one persistent launch.

    Scene s's photon p uses the Philox stream word pass | (stream_base + s) << 1,
    so its records equal those of trace(..., scene_id=stream_base + s) for the
    same key.  Returns
    (sinks, totals); sinks carry a scene array.  Capacities come from a pilot
    batch of every scene's first photons; an overflow reruns with exact sizes.
    """
    torch = _torch()
    L = load_library()
    counts = np.asarray(counts, dtype=np.int64)
    if len(counts) != table.n:
        raise ValueError("one photon count per scene")
    query = pass_ == abi.PASS_BACKWARD
    rng = abi.make_rng(key, pass_ | (int(stream_base) << 1))
    s = stream if stream is not None else _stream_ptr()
    total = int(counts.sum())
    if caps is None:
        pc = np.minimum(counts, np.maximum(np.ceil(counts * pilot_fraction).astype(np.int64), 256))
        if total > 8 * int(pc.sum()):
            _, pcnt = trace_batched(table, key, pass_, pc, False, None, block, refill, stream,
                                    caps=[int(r * pc.sum()) + 4096 for r in (4.0, 64.0, 1.0, 4.0)],
                                    stream_base=stream_base)
            frac = total / max(int(pc.sum()), 1)
            caps = [int(c * frac * 1.08) + 65536 for c in pcnt]
        else:
            caps = [int(r * total) + 4096 for r in (4.0, 64.0, 1.0, 4.0)]
    offsets = _dev_i64(np.concatenate([[0], np.cumsum(counts)]))
    work = torch.zeros(1, dtype=torch.int64, device=table.device)
    for attempt in range(2):
        sinks = [Sink.alloc(caps[t], query, tags, scenes=True) for t in range(abi.N_TYPES)]
        arr = ((abi.Queries if query else abi.Records) * 4)()
        for t in range(4):
            arr[t] = sinks[t].struct
        if attempt == 1:
            work.zero_()
            if counters is not None:
                counters.zero_()
        rc = L.alb_trace_batched(_p(table.dev), table.n, _p(offsets), C.byref(rng), pass_,
                                 None if query else C.cast(arr, C.c_void_p), C.cast(arr, C.c_void_p) if query else None,
                                 _p(work), block, refill, _p(counters), C.c_void_p(s))
        _check(rc, "alb_trace_batched")
        got = [int(x) for x in torch.cat([sk.tensors["count"] for sk in sinks]).cpu().tolist()]
        if all(got[t] <= caps[t] for t in range(4)):
            return sinks, got
        caps = [max(got[t], 1) for t in range(4)]
    raise AlbError("alb_trace_batched: capacity overflow persisted after resizing")


class BatchedTraceGraph:
    """One pass of the persistent batched trace captured in a HIP graph.

    The graph holds the reset of the work counter, the sink counts and the
    event counters (fill kernels) and the alb_trace_batched launch over
    table.dev, with sinks of fixed capacities.  replay() reruns it: after
    table.update(models) the same launch traces the new scenes (same scene
    count and photon counts).  It returns (sinks, counts) like trace_batched;
    the sinks belong to the graph and are overwritten by the next replay.
    counts above the capacities mean an overflow: replay() then raises
    MemoryError and the caller re-plans (a new graph with larger caps).
    """

    def __init__(self, table: DeviceSceneTable, key, pass_: int, counts, caps, tags: bool = False,
                 block: int = 256, refill: int = 4, stream_base: int = 0):
        torch = _torch()
        L = load_library()
        counts = np.asarray(counts, dtype=np.int64)
        if len(counts) != table.n:
            raise ValueError("one photon count per scene")
        self.table, self.pass_, self.caps = table, pass_, [int(c) for c in caps]
        query = pass_ == abi.PASS_BACKWARD
        self.rng = abi.make_rng(key, pass_ | (int(stream_base) << 1))
        self.offsets = _dev_i64(np.concatenate([[0], np.cumsum(counts)]))
        self.work = torch.zeros(1, dtype=torch.int64, device=table.device)
        self.counters = torch.zeros(abi.N_COUNTERS, dtype=torch.int64, device=table.device)
        self.sinks = [Sink.alloc(self.caps[t], query, tags, scenes=True) for t in range(abi.N_TYPES)]
        self.arr = ((abi.Queries if query else abi.Records) * 4)()
        for t in range(4):
            self.arr[t] = self.sinks[t].struct
        a = C.cast(self.arr, C.c_void_p)
        self._args = (_p(table.dev), table.n, _p(self.offsets), C.byref(self.rng), pass_,
                      None if query else a, a if query else None, _p(self.work), int(block), int(refill),
                      _p(self.counters))
        self._lib = L

        def body():
            self.work.zero_()
            for sk in self.sinks:
                sk.tensors["count"].zero_()
            self.counters.zero_()
            _check(L.alb_trace_batched(*self._args, C.c_void_p(torch.cuda.current_stream().cuda_stream)),
                   "alb_trace_batched")

        side = torch.cuda.Stream()  # default priority
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            body()  # one eager run before capture
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        self.graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(self.graph, capture_error_mode="relaxed"):
            body()

    def replay(self):
        torch = _torch()
        self.graph.replay()
        got = [int(x) for x in torch.cat([sk.tensors["count"] for sk in self.sinks]).cpu().tolist()]
        if any(got[t] > self.caps[t] for t in range(4)):
            raise MemoryError(f"alb_trace_batched: sink overflow in a captured launch ({got} > {self.caps})")
        return self.sinks, got


def build_map_batched(records: Sink, n: int, dims: int, n_scenes: int, stream=None) -> DeviceMap:
    """Sort n records of many scenes by (scene, Morton key) into one map whose
    tree holds every scene as a subtree (scene ranges and roots in m.index)."""
    torch = _torch()
    L = load_library()
    dev = records.tensors["x"].device
    m = DeviceMap(n=n, dims=dims)
    t = m.tensors
    for f in REC_F32:
        t[f] = torch.empty(max(n, 1), dtype=torch.float32, device=dev)
    t["perm"] = torch.empty(max(n, 1), dtype=torch.int32, device=dev)
    t["keys"] = torch.empty(max(n, 1), dtype=torch.int64, device=dev)
    t["meta"] = torch.zeros(8, dtype=torch.float64, device=dev)
    t["nodes"] = torch.empty((max(n - 1, 1), 16), dtype=torch.int32, device=dev)
    for f in ("first", "last", "root"):
        t[f] = torch.empty(n_scenes, dtype=torch.int32, device=dev)
    st = abi.Map()
    st.n = n
    st.dims = dims
    for f in ("x", "y", "z", "dx", "dy", "dz", "energy", "time", "perm", "keys", "meta", "nodes"):
        setattr(st, f, _p(t[f]))
    ix = abi.SceneIndex()
    ix.n_scenes = n_scenes
    ix.first, ix.last, ix.root = _p(t["first"]), _p(t["last"]), _p(t["root"])
    m.struct = st
    m.index = ix
    ws_bytes = C.c_size_t(0)
    _check(L.alb_map_workspace_size(n, C.byref(ws_bytes)), "alb_map_workspace_size")
    ws = torch.empty(max(ws_bytes.value, 1), dtype=torch.uint8, device=dev)
    s = stream if stream is not None else _stream_ptr()
    _check(L.alb_map_build_batched(C.byref(records.struct), n, dims, C.byref(st), C.byref(ix), _p(ws),
                                   ws_bytes.value, C.c_void_p(s)), "alb_map_build_batched")
    t["_ws"] = ws
    return m


def gather_batched(table: DeviceSceneTable, dmap: DeviceMap, ptype: int, queries: Sink, q_n: int, waveforms,
                   wf_offset, win0, win_len: int, counters=None, stream=None, events=None, nbr_out=None,
                   radius_out=None):
    """Batched k-NN gather of q_n queries (of many scenes) of one photon type.

    waveforms: flat float64 device tensor; scene s's type-t waveform starts at
    wf_offset[s] + t * n_bins_s.  win0: int32 device tensor of per-scene LDS
    window starts.  nbr_out / radius_out (optional, [q_n, k] / [q_n]) receive
    the rows in slot order; the slot order itself is returned.
    """
    torch = _torch()
    L = load_library()
    if q_n == 0:
        return None
    k = table.k
    s = stream if stream is not None else _stream_ptr()
    dev = waveforms.device
    b = C.c_size_t(0)
    _check(L.alb_gather_workspace_size(q_n, 1, C.byref(b)), "alb_gather_workspace_size")
    ws = torch.empty(b.value, dtype=torch.uint8, device=dev)
    order = torch.empty(q_n, dtype=torch.int32, device=dev)
    _check(L.alb_gather_order_batched(C.byref(dmap.struct), C.byref(dmap.index), C.byref(queries.struct), q_n,
                                      _p(order), _p(ws), b.value, C.c_void_p(s)), "alb_gather_order_batched")
    direct = nbr_out is not None
    chunk = q_n if direct else gather_chunk(q_n, k)
    nbr = nbr_out if direct else torch.empty((chunk, k), dtype=torch.int32, device=dev)
    rad = radius_out if direct else torch.empty(chunk, dtype=torch.float64, device=dev)
    knorm = torch.empty(chunk, dtype=torch.float64, device=dev)
    nb_max = int(table.n_bins.max())
    n_rep = BATCHED_REPLICAS
    rep = torch.empty(n_rep * table.n * nb_max, dtype=torch.float64, device=dev) if n_rep else None

    def ev():
        e = torch.cuda.Event(enable_timing=True)
        e.record(torch.cuda.current_stream())
        return e

    work = torch.empty(abi.KNN_WORK_WORDS, dtype=torch.int64, device=dev)  # cleared by each alb_knn_batched call
    for c0 in range(0, q_n, chunk):
        cn = min(chunk, q_n - c0)
        e0 = ev() if events is not None else None
        _check(L.alb_knn_batched(C.byref(dmap.struct), C.byref(dmap.index), ptype, k, C.byref(queries.struct),
                                 _p(order), c0, cn, _p(nbr), _p(rad), _p(knorm), _p(counters), _p(work),
                                 C.c_void_p(s)), "alb_knn_batched")
        e1 = ev() if events is not None else None
        _check(L.alb_contrib_batched(_p(table.dev), table.n, k, C.byref(dmap.struct), ptype, C.byref(queries.struct),
                                     _p(order), c0, cn, _p(nbr), _p(knorm), _p(waveforms), _p(wf_offset), _p(win0),
                                     int(win_len), _p(rep), n_rep, nb_max, C.c_void_p(s)), "alb_contrib_batched")
        if events is not None:
            e2 = ev()
            events.setdefault("knn", []).append((e0, e1))
            events.setdefault("contrib", []).append((e1, e2))
    return (ws, order, knorm, nbr, rad, rep)
