"""Many scenes per launch: the batched form of run_linear's forward trace ->
map build -> backward trace -> gather sequence.

A scene sweep (for example a Latin hypercube over depth, attenuation, phase
asymmetry, floor albedo and surface roughness) calls run_linear once per scene
in the reference (execution/linear.py:77-117 via
notebooks/run_hypercube_sampling.ipynb).  Here a group of scenes goes through
the HIP path together:

* one persistent trace launch per pass, lanes refilled from a global photon
  counter (block size and refill interval are tunable; see
  pyalbsim_amd.autotune);
* one map per photon type for the whole group, sorted by (scene, Morton key),
  whose radix tree holds every scene as a subtree;
* one k-NN and one contribution launch per type (per 2^22 queries), each
  query searching only its own scene's subtree.

Scene s of a group with scene ids ``ids`` uses the Philox stream word
pass | ids[s] << 1, so its waveform equals simulate(model_s, ..., scene_id =
ids[s]) up to FP64 summation order.
"""
from __future__ import annotations

import time
from dataclasses import dataclass, field
from typing import Optional, Sequence

import numpy as np

from .. import abi, runtime
from ..model import SceneModel
from ..photon_mapping.photon_type import PhotonType


@dataclass
class BatchedResult:
    models: list
    waveforms: object          # torch float64, flat: scene s at offsets[s], shape [4, n_bins_s]
    offsets: np.ndarray        # int64 [n_scenes]
    map_counts: list           # per type, records of all scenes
    query_counts: list
    scene_map_counts: np.ndarray   # [n_scenes, 4] records per scene and type
    counters_fwd: np.ndarray = None
    counters_bwd: np.ndarray = None
    timings: dict = field(default_factory=dict)
    maps: list = field(default_factory=list)
    queries: list = field(default_factory=list)
    state: dict = field(default_factory=dict)

    def waveform(self, s: int):
        """Device view [4, n_bins_s] of scene s."""
        nb = self.models[s].n_bins
        o = int(self.offsets[s])
        return self.waveforms[o:o + 4 * nb].view(4, nb)

    def gated(self, s: int, window=None):
        g0, gl = window or self.models[s].gate()
        return self.waveform(s)[:, g0:g0 + gl]


def simulate_batched(models: Sequence[SceneModel], n_forward, n_backward, key, *, scene_ids=None,
                     win_len: int = 1024, block=256, refill: int = 4, strict: bool = True,
                     keep: bool = False, tags: bool = False, events: Optional[dict] = None,
                     kernel_events: Optional[dict] = None, table: Optional[runtime.DeviceSceneTable] = None,
                     adjoint: bool = False, graphs: Optional[dict] = None) -> BatchedResult:
    """Forward trace, map build, backward trace and gather of many scenes.

    n_forward / n_backward: photons per scene (sequences or scalars).
    scene_ids: global scene numbers selecting the Philox streams (default
    0..n-1).  Only contiguous ids are supported by the trace kernel, so
    scene_ids must be range(base, base + n).
    win_len: LDS-privatised bins per scene, starting at model.gate(win_len)
    (widened to runtime.LDS_WINDOW_BATCHED bins; the window affects speed only).
    block / refill: launch parameters of the batched trace; block="auto" picks
    both with pyalbsim_amd.autotune on a pilot sample (cached per workload).
    strict: raise KeyError / IndexError, as the reference does per scene
    (core/simulation.py:134-153), when a scene has gather queries of a type
    whose map is empty or smaller than k; strict=False leaves those scenes'
    contributions of that type at zero.
    adjoint: keep what pyalbsim_amd.adjoint needs afterwards (record and
    query sinks with tags, the maps, each type's slot order, neighbour rows
    and kernel normalisations) in res.state.
    graphs: a dict the caller keeps across calls with the same scene count
    and photon counts (an optimisation loop, repeated sweeps): the scene
    table is then updated in place and both trace passes run as replays of
    HIP graphs captured on the first call (runtime.BatchedTraceGraph).  The
    sinks of a replayed pass belong to the graph: a kept result is valid
    until the next call with the same dict.
    """
    import torch

    models = list(models)
    runtime.check_k(models)
    n = len(models)
    nf = np.broadcast_to(np.asarray(n_forward, dtype=np.int64), (n,)).copy()
    nb = np.broadcast_to(np.asarray(n_backward, dtype=np.int64), (n,)).copy()
    base = 0
    if scene_ids is not None:
        ids = np.asarray(scene_ids, dtype=np.int64)
        base = int(ids[0])
        if not np.array_equal(ids, np.arange(base, base + n)):
            raise ValueError("scene_ids must be contiguous")
    t = {}

    def mark(name):
        if events is not None:
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(torch.cuda.current_stream())
            events.setdefault(name, []).append(ev)
        t[name] = time.perf_counter()

    mark("start")
    if adjoint:
        tags = keep = True
    if graphs is not None and table is None:
        tab = graphs.get("table")
        if tab is not None and tab.n == n and tab.adjoint == adjoint:
            tab.update(models)
        else:
            tab = runtime.DeviceSceneTable(models, adjoint=adjoint)
            graphs.clear()
            graphs["table"] = tab
    else:
        tab = table or runtime.DeviceSceneTable(models, adjoint=adjoint)
    if block == "auto":
        from ..autotune import tune_batched

        block, refill, _ = tune_batched(models, nf, key, table=tab)
    cf = torch.zeros(abi.N_COUNTERS, dtype=torch.int64, device="cuda")
    cb = torch.zeros(abi.N_COUNTERS, dtype=torch.int64, device="cuda")
    def trace_pass(pass_, counts, ctr):
        if graphs is None:
            return runtime.trace_batched(tab, key, pass_, counts, tags=tags, counters=ctr, block=block,
                                         refill=refill, stream_base=base)
        gk = (pass_, tuple(int(c) for c in counts), bool(tags), block, refill, base, tuple(key))
        g = graphs.get(gk)
        if g is not None:
            try:
                sinks, got = g.replay()
                ctr.copy_(g.counters)
                return sinks, got
            except MemoryError:
                pass  # more records than the graph's sinks hold: re-plan below
        sinks, got = runtime.trace_batched(tab, key, pass_, counts, tags=tags, counters=ctr, block=block,
                                           refill=refill, stream_base=base)
        caps = [int(c * 1.25) + 4096 for c in got]
        graphs[gk] = runtime.BatchedTraceGraph(tab, key, pass_, counts, caps, tags=tags, block=block,
                                               refill=refill, stream_base=base)
        return sinks, got

    recs, rcounts = trace_pass(abi.PASS_FORWARD, nf, cf)
    mark("forward")
    maps = [runtime.build_map_batched(recs[ti], rcounts[ti], runtime.MAP_DIMS[ti], n) for ti in range(abi.N_TYPES)]
    if not keep:
        del recs
    mark("build")
    qs, qcounts = trace_pass(abi.PASS_BACKWARD, nb, cb)
    mark("backward")
    sizes = np.array([4 * m.n_bins for m in models], dtype=np.int64)
    offsets = np.concatenate([[0], np.cumsum(sizes)[:-1]]).astype(np.int64)
    wf = torch.zeros(int(sizes.sum()), dtype=torch.float64, device="cuda")
    wf_off = torch.as_tensor(offsets, device="cuda")
    gates = [runtime.lds_window(m.n_bins, *m.gate(win_len), runtime.LDS_WINDOW_BATCHED) for m in models]
    wl = min(g[1] for g in gates)
    win0 = torch.as_tensor(np.array([g[0] for g in gates], dtype=np.int32), device="cuda")
    # records per scene and type, from the scene ranges of the maps
    per_scene = np.zeros((n, abi.N_TYPES), dtype=np.int64)
    for ti in range(abi.N_TYPES):
        if rcounts[ti]:
            f = maps[ti].tensors["first"].cpu().numpy().astype(np.int64)
            l_ = maps[ti].tensors["last"].cpu().numpy().astype(np.int64)
            per_scene[:, ti] = l_ - f + 1
    keep_alive = []
    rows = {}
    k = tab.k
    for ti in range(abi.N_TYPES):
        if qcounts[ti] == 0:
            continue
        qsc = torch.bincount(qs[ti].tensors["scene"][:qcounts[ti]].long(), minlength=n).cpu().numpy()
        bad = (qsc > 0) & (per_scene[:, ti] < k)
        if bad.any() and strict:
            s = int(np.nonzero(bad)[0][0])
            if per_scene[s, ti] == 0:
                raise KeyError(f"scene {base + s}: {PhotonType(ti + 1)}")
            raise IndexError(f"scene {base + s}: photon_mapping_k = {k} exceeds the {PhotonType(ti + 1).name} "
                             f"map size {per_scene[s, ti]}")
        if rcounts[ti] == 0:
            continue
        if adjoint:
            nbr = torch.empty((qcounts[ti], k), dtype=torch.int32, device="cuda")
            rad = torch.empty(qcounts[ti], dtype=torch.float64, device="cuda")
            kept = runtime.gather_batched(tab, maps[ti], ti, qs[ti], qcounts[ti], wf, wf_off, win0, wl, counters=cb,
                                          events=kernel_events, nbr_out=nbr, radius_out=rad)
            rows[ti] = {"order": kept[1], "knorm": kept[2], "nbr": nbr}
            keep_alive.append(kept)
        else:
            keep_alive.append(runtime.gather_batched(tab, maps[ti], ti, qs[ti], qcounts[ti], wf, wf_off, win0, wl,
                                                     counters=cb, events=kernel_events))
    mark("gather")
    ctr_f, ctr_b = cf.cpu().numpy().view(np.uint64), cb.cpu().numpy().view(np.uint64)
    if ctr_b[abi.CTR_OVERFLOW]:
        raise runtime.AlbError("gather node stack overflow")
    del keep_alive
    res = BatchedResult(models, wf, offsets, rcounts, qcounts, per_scene, ctr_f, ctr_b, timings=t)
    if keep:
        res.maps = maps
        res.queries = qs
    if adjoint:
        res.state = {"table": tab, "nf": nf, "nb": nb, "key": key, "base": base, "block": block, "refill": refill,
                     "records": recs, "rcounts": rcounts, "queries": qs, "qcounts": qcounts, "maps": maps,
                     "rows": rows, "wf_off": wf_off}
    return res
