"""run_parallel — the reference's multiprocess driver (src/alb_sim/execution/
parallel.py:186-284), same signature and return contract, on the GPU.

The reference splits the forward and backward batches over ``processes``
worker processes, merges the workers' photon storages into PhotonMapData and
sums the workers' waveforms.  Here the GPU's lanes are the workers: one
forward trace, one map build per type, one backward trace and one gather,
exactly as run_linear, and ``processes`` is ignored.  Returned:

  * waveform: {PhotonType: float32 [2*steps]} (the reference sums float32
    worker arrays);
  * photon_maps_data: {PhotonType: PhotonMapData} of the forward records in
    the map's Morton order (the reference's order is batch order; the records
    are the same set), with each record's first water and seafloor
    interaction points (NaN before the interaction);
  * sampled heat maps: SampledHeatmapsResult (water and seafloor energy,
    scatter-radius counts; float32 like the reference, summed in FP64) when
    simulation_config.heatmap.enabled, else None.

``seed`` seeds numpy.random.default_rng, whose two draws make the Philox key
(None: a fresh secrets-seeded generator, as the reference).  The photons of
all batches are distinct Philox streams.  The reference instead gives every
worker the same seed when one is passed (execution/parallel.py:66-68,
110-112, 208-210, 240-242), so its maps hold batches_forward identical
copies of one forward batch and its waveform sums batches_backward identical
backward batches; same_seed_batches=True reproduces that (with a seed): one
batch of each pass is traced, its records are repeated batches_forward
times before the map build, and the backward paths count batches_backward
times.
"""
from __future__ import annotations

import secrets
import time
from typing import Optional, Union

import numpy as np

from .. import abi
from ..config.run import RunConfig
from ..config.simulation import SimulationConfig
from ..model import build_scene
from ..photon_mapping.photon_map_data import PhotonMapData
from ..photon_mapping.photon_type import PhotonType
from .linear import photon_map_stats_table, simulate


def photon_maps_from_result(res) -> dict:
    """PhotonMapData per type from a simulate(..., keep=True) result (device -> host copy)."""
    out = {}
    for pt in PhotonType:
        t = pt.abi_index
        n = res.map_counts[t]
        m = res.maps[t].tensors
        pos = np.stack([m[f][:n].cpu().numpy() for f in ("x", "y", "z")], 1).astype(np.float32)
        dirs = np.stack([m[f][:n].cpu().numpy() for f in ("dx", "dy", "dz")], 1).astype(np.float32)
        if m.get("fwi_x") is not None:
            fwi = np.stack([m[f][:n].cpu().numpy() for f in ("fwi_x", "fwi_y", "fwi_z")], 1).astype(np.float32)
            sfi = np.stack([m[f][:n].cpu().numpy() for f in ("sfi_x", "sfi_y", "sfi_z")], 1).astype(np.float32)
        else:
            fwi = np.full((n, 3), np.nan, dtype=np.float32)
            sfi = fwi.copy()
        out[pt] = PhotonMapData(positions=pos, directions=dirs, energies=m["energy"][:n].cpu().numpy().copy(),
                                times=m["time"][:n].cpu().numpy().copy(), first_water_interaction=fwi,
                                seafloor_interaction=sfi)
    return out


def run_parallel(simulation_config: SimulationConfig, run_config: RunConfig,
                 seed: Union[int, None] = None, *, verbose: bool = True, same_seed_batches: bool = False):
    """(waveform, photon_maps_data, sampled heat maps or None) for one scene, computed on the GPU."""
    rng = np.random.default_rng(secrets.randbits(128) if seed is None else seed)
    key = abi.key_from_generator(rng)
    model = build_scene(simulation_config)
    n_f = run_config.photons_per_batch_forward * run_config.batches_forward
    n_b = run_config.photons_per_batch_backward * run_config.batches_backward
    copies = (1, 1)
    if same_seed_batches and seed is not None:
        n_f, n_b = run_config.photons_per_batch_forward, run_config.photons_per_batch_backward
        copies = (run_config.batches_forward, run_config.batches_backward)
    if verbose:
        print("Starting forwards pass")
    t0 = time.perf_counter()
    hc = simulation_config.heatmap
    res = simulate(model, n_f, n_b, key, keep=True, points=True, heatmaps=hc if hc.enabled else None,
                   copies=copies)
    maps = photon_maps_from_result(res)
    if verbose:
        print("\n" + photon_map_stats_table(res.map_counts) + "\n")
        print(f"Finished {n_f * copies[0]} + {n_b * copies[1]} photon paths in {time.perf_counter() - t0:.2f} s")
    wf = res.waveforms.cpu().numpy()
    waveform = {pt: wf[pt.abi_index].astype(np.float32) for pt in PhotonType}
    heat = None
    if res.heatmaps is not None:
        from ..utils.heatmap_result import SampledHeatmapsResult

        h = res.heatmaps
        heat = SampledHeatmapsResult(water_heatmap=h["water"].astype(np.float32),
                                     seafloor_heatmap=h["floor"].astype(np.float32),
                                     scatter_radius_heatmap=h["radius"].astype(np.float32))
    return waveform, maps, heat
