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


@dataclass
class SweepResult:
    groups: list          # (first, end) scene ranges run as one batched group each
    waveforms: list       # per scene: float64 numpy [4, win_len] over model.gate(win_len)
    windows: list         # per scene: (first bin, length) of its waveform
    seconds: float


def estimate_group_bytes(models, n_forward, n_backward, key, pilot_fraction: float = 1 / 100,
                         bytes_per_record: float = 240.0, bytes_per_query: float = 80.0) -> np.ndarray:
    """Device bytes each scene needs inside a batched group: pilot forward and
    backward traces of pilot_fraction of every scene's photons give its map
    records per forward photon and its gather queries per backward photon;
    records cost sinks + map + radix tree + build workspace
    (bytes_per_record), queries their sinks, order and neighbour rows
    (bytes_per_query).  Backward paths keep scattering until their time
    budget ends, so a scene's queries per photon differ from its records per
    photon and are counted, not inferred."""
    import torch

    from . import abi, runtime

    nf = np.asarray(n_forward, dtype=np.int64)
    nb = np.asarray(n_backward, dtype=np.int64)
    tab = runtime.DeviceSceneTable(list(models))

    def per_photon(n, pass_):
        pc = np.maximum(np.ceil(n * pilot_fraction).astype(np.int64), np.minimum(n, 256))
        sinks, got = runtime.trace_batched(tab, key, pass_, pc)
        cnt = np.zeros(len(n))
        for t in range(abi.N_TYPES):
            if got[t]:
                cnt += torch.bincount(sinks[t].tensors["scene"][:got[t]].long(), minlength=len(n)).cpu().numpy()
        return cnt / np.maximum(pc, 1)

    rec = per_photon(nf, abi.PASS_FORWARD) * nf
    qry = per_photon(nb, abi.PASS_BACKWARD) * nb
    return rec * bytes_per_record + qry * bytes_per_query


def free_device_bytes() -> int:
    """Device memory free for new allocations, including the blocks torch's
    caching allocator holds but no tensor uses (released first)."""
    import torch

    torch.cuda.empty_cache()
    free, _ = torch.cuda.mem_get_info()
    return int(free)


def plan_groups(scene_bytes, budget: float) -> list:
    """Contiguous scene ranges whose summed bytes stay within budget (a scene
    larger than the budget forms a group of its own)."""
    groups, start, acc = [], 0, 0.0
    for s, b in enumerate(scene_bytes):
        if s > start and acc + b > budget:
            groups.append((start, s))
            start, acc = s, 0.0
        acc += float(b)
    if len(scene_bytes):
        groups.append((start, len(scene_bytes)))
    return groups


def run_sweep(scenes, key, budget_bytes: float = None, win_len: int = 1000, margin: float = 1.3, log=None,
              **batched_kw) -> SweepResult:
    """All scenes of a sweep in the batched form, in contiguous groups whose
    photon maps fit the device: a full-size sweep of 1024 scenes at 10^U[5,7]
    photons holds ~1e11 map records, far beyond one GPU's memory, so the
    scenes are packed into groups by a pilot estimate (estimate_group_bytes,
    times margin) against budget_bytes (default: half the free device
    memory).  Scene s keeps Philox stream id s in every grouping, so its
    waveform does not depend on the grouping.  Returns each scene's waveform
    over its gate window (model.gate(win_len)); log, if given, receives a
    line per finished group."""
    import time as _time

    import torch

    from .execution.batched import simulate_batched

    models = [x.model for x in scenes]
    nf = np.array([x.n_forward for x in scenes], dtype=np.int64)
    nb = np.array([x.n_backward for x in scenes], dtype=np.int64)
    t0 = _time.perf_counter()
    if budget_bytes is None:
        budget_bytes = 0.5 * free_device_bytes()
    est = estimate_group_bytes(models, nf, nb, key) * margin
    groups = plan_groups(est, budget_bytes)
    wfs, wins = [], []
    for a, b in groups:
        res = simulate_batched(models[a:b], nf[a:b], nb[a:b], key, scene_ids=range(a, b), win_len=win_len,
                               **batched_kw)
        for s in range(b - a):
            w = models[a + s].gate(win_len)
            wins.append(w)
            wfs.append(res.gated(s, w).cpu().numpy())
        del res
        if log is not None:
            log(f"sweep group {a}-{b} of {len(models)} scenes done")
    torch.cuda.synchronize()
    return SweepResult(groups, wfs, wins, _time.perf_counter() - t0)
