"""Interaction points and sampled heat maps on the GPU vs the oracle.

* forward records' first water / seafloor interaction points match the
  oracle's photon by photon (tags; the trajectory tolerance of
  test_gpu_trace_parity, NaN where the oracle has NaN);
* the forward scatter-radius counts match the oracle's on the same photons
  (>= 99.5 % of the counts in the same bin: a hit point differing by
  1e-11 relative can change a bin);
* gathered water / seafloor heat maps equal the oracle's sequential FP64
  re-sum over the same map (with its interaction points) and queries to
  1e-9 of the peak;
* run_parallel returns SampledHeatmapsResult when the configuration enables
  heat maps, and interaction points in its PhotonMapData.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from pyalbsim_amd import abi, runtime
from pyalbsim_amd.config import HeatmapConfig, RunConfig
from pyalbsim_amd.execution.linear import simulate
from pyalbsim_amd.execution.parallel import run_parallel
from pyalbsim_amd.model import build_scene
from tests.scenes import layered_scene, single_shot

pytestmark = pytest.mark.gpu

F = ("x", "y", "z", "dx", "dy", "dz", "energy", "time")
P = abi.INTERACTION_FIELDS
HEAT = {"bins": 40, "water_extent": 1.0, "seafloor_extent": 2.0}


@pytest.mark.parametrize("scene", ["single_shot", "layered"])
def test_interaction_points_and_radius_counts_match_oracle(gpu, scene):
    model = build_scene(single_shot(b=0.8) if scene == "single_shot" else layered_scene())
    ds = runtime.DeviceScene(model)
    ds.enable_heatmaps(HEAT["bins"], HEAT["water_extent"], HEAT["seafloor_extent"])
    n = 3000
    sinks, counts = runtime.trace(ds, (3, 4), abi.PASS_FORWARD, n, tags=True, points=True)
    heat = dict(HEAT)
    recs, _, _ = oracle.trace(model, (3, 4), abi.PASS_FORWARD, 0, n, points=True, heat=heat)
    bad = total = 0
    for t in range(4):
        g = {f: sinks[t].tensors[f][:counts[t]].cpu().numpy() for f in (*F, *P, "tag")}
        c = recs[t]
        gi = {int(v): i for i, v in enumerate(g["tag"])}
        for j, tag in enumerate(c["tag"]):
            i = gi.get(int(tag))
            total += 1
            if i is None:
                bad += 1
                continue
            for f in P:
                a, b = float(g[f][i]), float(c[f][j])
                if np.isnan(b):
                    ok = np.isnan(a)
                else:
                    ok = abs(a - b) <= 1e-5 * max(abs(b), 1.0)
                if not ok:
                    bad += 1
                    break
    assert total > 2000 and bad <= 0.005 * total, (bad, total)
    got = ds.heatmaps()["radius"].astype(np.int64)
    ref = heat["radius"].astype(np.int64)
    assert ref.sum() > 500
    assert np.abs(got - ref).sum() <= 0.005 * ref.sum() + 2


def test_gathered_heatmaps_match_oracle(gpu):
    torch = gpu
    model = build_scene(single_shot(b=0.8))
    ds = runtime.DeviceScene(model)
    ds.enable_heatmaps(HEAT["bins"], HEAT["water_extent"], HEAT["seafloor_extent"])
    recs, rc = runtime.trace(ds, (5, 6), abi.PASS_FORWARD, 20000, points=True)
    maps = [runtime.build_map(recs[t], rc[t], runtime.MAP_DIMS[t]) for t in range(4)]
    qs, qc = runtime.trace(ds, (5, 6), abi.PASS_BACKWARD, 1000)
    for t in range(4):
        nq = min(qc[t], 500)
        if nq == 0 or rc[t] < model.k:
            continue
        ds.clear_heatmaps()
        wf = torch.zeros(model.n_bins, dtype=torch.float64, device="cuda")
        runtime.gather(ds, maps[t], t, qs[t], nq, wf, model.gate())
        torch.cuda.synchronize()
        got = ds.heatmaps()
        mp = {f: maps[t].tensors[f][:rc[t]].cpu().numpy() for f in (*F, *P)}
        qv = {f: qs[t].tensors[f][:nq].cpu().numpy() for f in F}
        heat = dict(HEAT)
        oracle.gather(model, t, mp, qv, heat=heat)
        for key in ("water", "floor"):
            ref = heat[key]
            peak = np.max(np.abs(ref))
            if peak == 0:
                assert np.max(np.abs(got[key])) == 0, (t, key)
            else:
                assert np.max(np.abs(got[key] - ref)) <= 1e-9 * peak, (t, key)
        assert heat["water"].sum() > 0


def test_run_parallel_heatmaps_and_points(gpu):
    cfg = dataclasses.replace(single_shot(b=0.8), heatmap=HeatmapConfig(enabled=True, bins=50))
    rc = RunConfig(photons_per_batch_forward=20000, batches_forward=1, photons_per_batch_backward=2000,
                   batches_backward=1)
    wf, maps, heat = run_parallel(cfg, rc, seed=3, verbose=False)
    assert heat is not None
    for h in (heat.water_heatmap, heat.seafloor_heatmap, heat.scatter_radius_heatmap):
        assert h.shape == (50, 50) and h.dtype == np.float32 and np.all(np.isfinite(h)) and h.sum() > 0
    from pyalbsim_amd.photon_mapping.photon_type import PhotonType

    b = maps[PhotonType.BOTTOM_REFLECTION]
    assert np.array_equal(b.seafloor_interaction, b.positions)  # a floor record's own hit point
    assert not np.any(np.isnan(b.first_water_interaction))
    s = maps[PhotonType.SURFACE_REFLECTION]
    assert np.array_equal(s.first_water_interaction, s.positions)
    assert np.all(np.isnan(s.seafloor_interaction))
    cfg_off = dataclasses.replace(cfg, heatmap=HeatmapConfig(enabled=False))
    wf2, _, heat2 = run_parallel(cfg_off, rc, seed=3, verbose=False)
    assert heat2 is None
    for pt in PhotonType:  # heat maps do not change the waveform (FP64 atomics: summation order only)
        assert np.allclose(wf[pt], wf2[pt], rtol=1e-6, atol=1e-6 * float(np.max(np.abs(wf2[pt]))))


def test_simulate_with_heatmaps_keeps_waveforms(gpu):
    model = build_scene(single_shot())
    a = simulate(model, 20000, 2000, (8, 9)).waveforms.cpu().numpy()
    res = simulate(model, 20000, 2000, (8, 9), heatmaps=HeatmapConfig(enabled=True, bins=30))
    b = res.waveforms.cpu().numpy()
    for t in range(4):
        peak = float(np.max(np.abs(a[t])))
        assert np.max(np.abs(a[t] - b[t])) <= 1e-9 * peak + 1e-300
    assert res.heatmaps["water"].shape == (30, 30)
