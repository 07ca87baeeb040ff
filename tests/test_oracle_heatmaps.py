"""Interaction points and sampled heat maps in the oracle (CPU).

* Interaction points follow the reference's order of operations
  (core/simulation.py:393, 521: set before the record is stored): every
  forward record carries the photon's surface entry point, which is the
  position of its SURFACE_REFLECTION record; a BOTTOM record carries its own
  position as seafloor interaction; any later record carries the position of
  the photon's latest BOTTOM record, earlier ones NaN.
* Gathered heat maps equal accumulate_to_heatmap (math/accumulate_heatmap.py:
  6-61) restated with NumPy (astype(int64) truncation, NaN filter, np.add.at)
  over the per-pair energies of the BOTTOM gather, recomputed from the
  neighbour sets.
* The scatter-radius counts equal accumulate_scatter_radius_heatmap
  (math/accumulate_heatmap.py:64-124) restated with NumPy over the forward
  BOTTOM records (float32 hit points in place of the FP64 ones, so a few
  counts may move to a neighbouring bin).
"""
import numpy as np

import oracle
from pyalbsim_amd import abi
from pyalbsim_amd.footprint import seafloor_intersection, water_surface_intersection
from pyalbsim_amd.model import build_scene
from tests.scenes import single_shot

F = ("x", "y", "z", "dx", "dy", "dz", "energy", "time")
HEAT = {"bins": 40, "water_extent": 1.0, "seafloor_extent": 2.0}


def _trace(model, n):
    heat = dict(HEAT)
    recs, _, _ = oracle.trace(model, (3, 4), abi.PASS_FORWARD, 0, n, points=True, heat=heat)
    return recs, heat


def test_interaction_points_follow_the_photon():
    model = build_scene(single_shot(b=0.8))
    recs, _ = _trace(model, 600)
    sr = recs[abi.SURFACE_REFLECTION]
    entry = {int(t >> 16): (sr["x"][i], sr["y"][i], sr["z"][i]) for i, t in enumerate(sr["tag"])}
    floors = {}
    bt = recs[abi.BOTTOM_REFLECTION]
    for i, t in enumerate(bt["tag"]):
        floors.setdefault(int(t >> 16), []).append((int(t & 0xFFFF), (bt["x"][i], bt["y"][i], bt["z"][i])))
    checked = after_floor = 0
    for ti in range(4):
        r = recs[ti]
        for i, t in enumerate(r["tag"]):
            p, ev = int(t >> 16), int(t & 0xFFFF)
            assert (r["fwi_x"][i], r["fwi_y"][i], r["fwi_z"][i]) == entry[p]
            prior = [pos for e, pos in floors.get(p, []) if e <= ev]
            if prior:
                assert (r["sfi_x"][i], r["sfi_y"][i], r["sfi_z"][i]) == prior[-1]
                after_floor += ti != abi.BOTTOM_REFLECTION
            else:
                assert np.isnan(r["sfi_x"][i]) and np.isnan(r["sfi_z"][i])
            checked += 1
    assert checked > 1000 and after_floor > 20


def _numpy_heatmap(px, pz, e, center, extent, bins):
    """accumulate_to_heatmap restated with NumPy."""
    h = np.zeros((bins, bins))
    valid = ~np.isnan(px)
    xc = px[valid].astype(np.float64) - center[0]
    zc = pz[valid].astype(np.float64) - center[2]
    bs = 2 * extent / bins
    xi = ((xc + extent) / bs).astype(np.int64)
    zi = ((zc + extent) / bs).astype(np.int64)
    ok = (xi >= 0) & (xi < bins) & (zi >= 0) & (zi < bins)
    np.add.at(h.ravel(), xi[ok] * bins + zi[ok], e[valid][ok])
    return h


def test_gathered_heatmaps_match_numpy_restatement():
    model = build_scene(single_shot(b=0.8))
    recs, _ = _trace(model, 4000)
    qs, _, _ = oracle.trace(model, (3, 4), abi.PASS_BACKWARD, 0, 300)
    t = abi.BOTTOM_REFLECTION
    mp, q = recs[t], qs[t]
    heat = dict(HEAT)
    wf, nbr, rad = oracle.gather(model, t, mp, q, want_nbrs=True, heat=heat)
    # per-pair energies of the BOTTOM gather (core/simulation.py:155-195)
    knorm = 1.0 / (np.pi * rad ** 2)
    mult = (model.floor_albedo / np.pi) * np.maximum(0.0, -q["dy"])
    e = ((q["energy"][:, None] * mp["energy"][nbr]).astype(np.float64) * knorm[:, None] * mult[:, None]).ravel()
    idx = nbr.ravel()
    wc, fc = water_surface_intersection(model), seafloor_intersection(model)
    ref_w = _numpy_heatmap(mp["fwi_x"][idx], mp["fwi_z"][idx], e, wc, HEAT["water_extent"], HEAT["bins"])
    ref_f = _numpy_heatmap(mp["sfi_x"][idx], mp["sfi_z"][idx], e, fc, HEAT["seafloor_extent"], HEAT["bins"])
    for got, ref in ((heat["water"], ref_w), (heat["floor"], ref_f)):
        assert ref.sum() > 0
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(ref)
    # a wider extent takes every pair: each map then holds the waveform's total
    wide = {"bins": 8, "water_extent": 1e3, "seafloor_extent": 1e3}
    oracle.gather(model, t, mp, q, heat=wide)
    assert abs(wide["water"].sum() - wf.sum()) <= 1e-9 * wf.sum()
    assert abs(wide["floor"].sum() - wf.sum()) <= 1e-9 * wf.sum()


def test_radius_counts_match_numpy_restatement():
    model = build_scene(single_shot(b=0.8))
    recs, heat = _trace(model, 4000)
    bt = recs[abi.BOTTOM_REFLECTION]
    bins = HEAT["bins"]
    wc, fc = water_surface_intersection(model), seafloor_intersection(model)
    rs = np.sqrt((bt["fwi_x"] - wc[0]) ** 2 + (bt["fwi_z"] - wc[2]) ** 2)
    rf = np.sqrt((bt["x"].astype(np.float64) - fc[0]) ** 2 + (bt["z"].astype(np.float64) - fc[2]) ** 2)
    xi = ((rs / HEAT["water_extent"]) * bins).astype(np.int64)
    yi = ((rf / HEAT["seafloor_extent"]) * bins).astype(np.int64)
    ok = (xi >= 0) & (xi < bins) & (yi >= 0) & (yi < bins)
    ref = np.zeros((bins, bins))
    np.add.at(ref.ravel(), xi[ok] * bins + yi[ok], 1)
    got = heat["radius"].astype(np.float64)
    assert got.sum() > 1000
    assert abs(got.sum() - ref.sum()) <= 2
    assert np.abs(got - ref).sum() <= 0.002 * ref.sum() + 4
