"""run_linear — the drop-in entry point (src/alb_sim/execution/linear.py:77-117).

Same signature and return contract as the reference: a dict keyed by the four
PhotonType members (enum order) holding waveforms of 2*steps bins of width
1/sample_rate.  Underneath, the forward pass, the photon-map build and the
backward gather run as HIP kernels on the current GPU.

Differences from the reference, by design:
  * photons are independent Philox4x32-10 streams keyed from ``rng`` (two
    draws of ``rng.integers``), so the result does not depend on
    photons_per_batch_* — only on the photon totals and the key;
  * waveform bins are accumulated in FP64 (returned as float64 unless
    ``as_float32``);
  * heat maps are not computed (the reference computes and discards them).
"""
from __future__ import annotations

import secrets
import time
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

from .. import abi, runtime
from ..config.run import RunConfig
from ..config.simulation import SimulationConfig
from ..model import SceneModel, build_scene
from ..photon_mapping.photon_type import PhotonType


@dataclass
class SimulationResult:
    """Device-side outcome of one scene: waveforms plus the intermediate maps."""

    model: SceneModel
    waveforms: object  # torch float64 [4, n_bins] on the device
    map_counts: list
    query_counts: list
    counters_fwd: np.ndarray
    counters_bwd: np.ndarray
    maps: list = field(default_factory=list)
    queries: list = field(default_factory=list)
    timings: dict = field(default_factory=dict)
    heatmaps: Optional[dict] = None  # water / floor energy and radius counts [bins, bins], when computed


def simulate(model: SceneModel, n_forward: int, n_backward: int, key, *, window=None, keep=False,
             tags: bool = False, dscene: Optional[runtime.DeviceScene] = None, sync_timing: bool = False,
             strict: bool = True, events: Optional[dict] = None,
             kernel_events: Optional[dict] = None, scene_id: int = 0, heatmaps=None,
             points: bool = False, replay_forward=None, replay_backward=None,
             copies=(1, 1)) -> SimulationResult:
    """Forward trace -> map build -> backward trace -> gather, all on the GPU.

    events: optional dict; HIP events are recorded on the launch stream at the
    phase boundaries start / forward / build / backward / gather.
    kernel_events: optional dict of (start, stop) event pairs around every
    k-NN ("knn") and contribution ("contrib") launch.
    strict: raise KeyError when a photon type has gather queries but an empty
    map and IndexError when k exceeds a map's size, as the reference does
    (core/simulation.py:134-153); strict=False gives such types zero signal.
    scene_id: Philox stream word pass | scene_id << 1 (scene scene_id of a
    batched run uses the same streams).
    heatmaps: a HeatmapConfig (or None): also compute the sampled heat maps
    (core/simulation.py:51-69, 197-214, 523-537) into res.heatmaps.
    points: keep the records' first water / seafloor interaction points in
    the maps (implied by heatmaps).
    replay_forward / replay_backward: optional host-supplied uniforms, FP64
    arrays [n_forward, L] / [n_backward, L] (photon slot i draws row i in
    order) in place of the Philox streams of that pass.
    copies: (forward, backward) repetitions of the traced photons: the maps
    hold every forward record `forward` times, and the waveforms and heat
    maps count every backward path `backward` times, as identical batches
    would (the reference's run_parallel with a seed gives every batch the
    same seed, execution/parallel.py:66-68, 110-112).
    """
    import torch

    runtime.check_k([model])
    ds = dscene or runtime.DeviceScene(model)
    if heatmaps is not None and heatmaps.enabled:
        if ds.heat is None or ds.heat["bins"] != int(heatmaps.bins):
            ds.enable_heatmaps(heatmaps.bins, heatmaps.water_extent, heatmaps.seafloor_extent)
        ds.clear_heatmaps()
        points = True
    elif ds.heat is not None:
        ds = runtime.DeviceScene(model)  # a scene without heat maps for this run
    win = window or model.gate()
    t = {}

    def mark(name):
        if events is not None:
            ev = torch.cuda.Event(enable_timing=True)
            ev.record(torch.cuda.current_stream())
            events.setdefault(name, []).append(ev)
        if sync_timing:
            torch.cuda.synchronize()
        t[name] = time.perf_counter()

    mark("start")
    cf = torch.zeros(abi.N_COUNTERS, dtype=torch.int64, device="cuda")
    cb = torch.zeros(abi.N_COUNTERS, dtype=torch.int64, device="cuda")
    rep_f = _device_replay(replay_forward, n_forward)
    rep_b = _device_replay(replay_backward, n_backward)
    cp_f, cp_b = (int(copies[0]), int(copies[1]))
    if cp_f < 1 or cp_b < 1:
        raise ValueError("copies must be positive")
    recs, rcounts = runtime.trace(ds, key, abi.PASS_FORWARD, n_forward, tags=tags, counters=cf, scene_id=scene_id,
                                  points=points, replay=rep_f)
    if cp_f > 1:
        recs, rcounts = _repeat_records(recs, rcounts, cp_f)
        if ds.heat is not None:  # every copy of a batch counts its floor hits again
            ds.heat["radius"].mul_(cp_f)
    mark("forward")
    maps = [runtime.build_map(recs[ti], rcounts[ti], runtime.MAP_DIMS[ti]) for ti in range(abi.N_TYPES)]
    if not keep:
        del recs
    mark("build")
    qs, qcounts = runtime.trace(ds, key, abi.PASS_BACKWARD, n_backward, tags=tags, counters=cb, scene_id=scene_id,
                                replay=rep_b)
    mark("backward")
    wf = torch.zeros((abi.N_TYPES, model.n_bins), dtype=torch.float64, device="cuda")
    keep_alive = []
    for ti in range(abi.N_TYPES):
        if qcounts[ti] == 0:
            continue
        if rcounts[ti] == 0 or model.k > rcounts[ti]:
            if not strict:
                continue
            if rcounts[ti] == 0:
                raise KeyError(PhotonType(ti + 1))
            raise IndexError(f"photon_mapping_k = {model.k} exceeds the {PhotonType(ti + 1).name} map size "
                             f"{rcounts[ti]}")
        keep_alive.append(runtime.gather(ds, maps[ti], ti, qs[ti], qcounts[ti], wf[ti], win, counters=cb,
                                         events=kernel_events))
    mark("gather")
    if cp_b > 1:
        wf.mul_(cp_b)
        if ds.heat is not None:
            ds.heat["water"].mul_(cp_b)
            ds.heat["floor"].mul_(cp_b)
    res = SimulationResult(model, wf, rcounts, qcounts, None, None, timings=t)
    ctr_f, ctr_b = cf.cpu().numpy().view(np.uint64), cb.cpu().numpy().view(np.uint64)
    if ctr_b[abi.CTR_OVERFLOW]:
        raise runtime.AlbError("gather node stack overflow")
    res.counters_fwd, res.counters_bwd = ctr_f, ctr_b
    del keep_alive  # the counter read-back above synchronised the stream
    if ds.heat is not None:
        res.heatmaps = ds.heatmaps()
    if keep:
        res.maps = maps
        res.queries = qs
    return res


def _repeat_records(sinks, counts, c: int):
    """Record sinks holding every record of sinks c times (copy j after copy j - 1)."""
    out, nc = [], []
    for sk, n in zip(sinks, counts):
        t = sk.tensors
        r = runtime.Sink.alloc(n * c, False, t.get("tag") is not None, device=t["x"].device,
                               points=t.get("fwi_x") is not None)
        for f, v in t.items():
            if f == "count" or v is None:
                continue
            r.tensors[f][:n * c].copy_(v[:n].repeat(c))
        r.tensors["count"].fill_(n * c)
        out.append(r)
        nc.append(n * c)
    return out, nc


def photon_map_stats_table(counts) -> str:
    from tabulate import tabulate

    rows = [[pt.name, counts[pt.abi_index], f"{counts[pt.abi_index] * 32 / 2 ** 20:.1f}"] for pt in PhotonType]
    return tabulate(rows, headers=["Photon type", "Records", "Device MiB (32 B/record)"], tablefmt="github")


def _device_replay(arr, n):
    if arr is None:
        return None
    import torch

    a = np.ascontiguousarray(arr, dtype=np.float64)
    if a.ndim != 2 or a.shape[0] != n:
        raise ValueError(f"replay uniforms must be [{n}, L]")
    return torch.as_tensor(a, device="cuda")


def run_linear(simulation_config: SimulationConfig, run_config: RunConfig,
               rng: Optional[np.random.Generator] = None, *, as_float32: bool = False,
               window=None, gate=None, rng_mode: str = "philox", replay_uniforms=None,
               device_tensors: bool = False, verbose: bool = True) -> dict:
    """Full forward + backward simulation of one scene on the GPU.

    Same contract as the reference run_linear (execution/linear.py:77-117):
    returns {PhotonType: waveform}, consumes ``rng`` (a fresh secrets-seeded
    generator when None).  Options beyond the reference's: gate (= window):
    (first bin, length) of the LDS-privatised waveform window; rng_mode
    "replay" with replay_uniforms = (forward [N_f, L], backward [N_b, L]) host
    FP64 uniforms per photon slot instead of the Philox streams (rng is then
    not consumed); device_tensors: return torch float64 device tensors.
    """
    if rng_mode not in ("philox", "replay"):
        raise ValueError("rng_mode must be 'philox' or 'replay'")
    if rng_mode == "replay":
        if replay_uniforms is None or len(replay_uniforms) != 2:
            raise ValueError("rng_mode='replay' needs replay_uniforms = (forward, backward)")
        key = (0, 0)
    else:
        if rng is None:
            rng = np.random.default_rng(secrets.randbits(128))
        key = abi.key_from_generator(rng)
    window = gate if gate is not None else window
    model = build_scene(simulation_config)
    n_f = run_config.photons_per_batch_forward * run_config.batches_forward
    n_b = run_config.photons_per_batch_backward * run_config.batches_backward
    if verbose:
        print("Starting forwards pass")
    t0 = time.perf_counter()
    rf, rb = replay_uniforms if rng_mode == "replay" else (None, None)
    res = simulate(model, n_f, n_b, key, window=window, replay_forward=rf, replay_backward=rb)
    if device_tensors:
        return {pt: res.waveforms[pt.abi_index] for pt in PhotonType}
    wf = res.waveforms.cpu().numpy()
    if verbose:
        print(f"Finished forwards and backwards pass: {n_f} + {n_b} photon paths in "
              f"{time.perf_counter() - t0:.2f} s")
        print("\n" + photon_map_stats_table(res.map_counts) + "\n")
    dt = np.float32 if as_float32 else np.float64
    return {pt: wf[pt.abi_index].astype(dt) for pt in PhotonType}
