"""Tilted seafloor (an extension of the reference's flat floor) in the oracle.

* The generalised forms reduce to the flat ones exactly: a flat scene marked
  tilted (normal +y, offset y_sf, frame = the axes) traces bit for bit like
  the flat scene, so the tilted code path restates the flat handlers with y
  replaced by the plane's signed distance.
* Floor records lie on the tilted plane (float32 positions), spread in depth
  over the footprint as the slope implies, and the plane passes through the
  beam's floor point.
* The BOTTOM gather multiplier equals (albedo / pi) max(0, n . (-d)) computed
  with NumPy.
"""
import dataclasses

import numpy as np
import pytest

import oracle
from pyalbsim_amd import abi
from pyalbsim_amd.model import beam_floor_point, build_scene
from tests.scenes import single_shot, tilted_shot


@pytest.mark.parametrize("pass_", [abi.PASS_FORWARD, abi.PASS_BACKWARD])
def test_flat_scene_through_the_tilted_path_is_unchanged(pass_):
    flat = build_scene(single_shot(b=0.6))
    forced = dataclasses.replace(flat, tilted=True)
    a, ca, _ = oracle.trace(flat, (4, 5), pass_, 0, 300)
    b, cb, _ = oracle.trace(forced, (4, 5), pass_, 0, 300)
    assert sum(len(a[t]["x"]) for t in range(4)) > 300
    for t in range(4):
        for f in a[t]:
            assert np.array_equal(a[t][f], b[t][f]), (t, f)
    assert np.array_equal(ca, cb)


def test_floor_records_lie_on_the_tilted_plane():
    m = build_scene(tilted_shot(slope=12.0, azimuth=70.0, b=0.1))
    assert m.tilted and not m.layers
    p0 = beam_floor_point(m.laser_dir, m.y_ws, m.depth, m.n_water)
    assert abs(float(np.dot(m.floor_normal, p0)) - m.floor_offset) < 1e-9
    recs, _, _ = oracle.trace(m, (6, 7), abi.PASS_FORWARD, 0, 3000)
    bt = recs[abi.BOTTOM_REFLECTION]
    pts = np.stack([bt["x"], bt["y"], bt["z"]], 1).astype(np.float64)
    dist = pts @ m.floor_normal - m.floor_offset
    assert len(pts) > 500
    assert np.max(np.abs(dist)) < 2e-4  # float32 positions near 405 m
    # the floor's depth varies across the footprint as the slope implies
    horiz = pts[:, [0, 2]] - p0[[0, 2]]
    lean = np.array([m.floor_normal[0], m.floor_normal[2]]) / np.hypot(m.floor_normal[0], m.floor_normal[2])
    dy = pts[:, 1] - p0[1]
    fit = np.polyfit(horiz @ lean, dy, 1)[0]
    assert fit == pytest.approx(-np.tan(np.radians(12.0)), rel=1e-3)


def test_bottom_multiplier_uses_the_normal():
    m = build_scene(tilted_shot(slope=15.0, azimuth=200.0))
    rs = np.random.default_rng(0)
    mp = {f: rs.uniform(-1, 1, 500).astype(np.float32) for f in ("x", "z", "dx", "dy", "dz")}
    mp["y"] = np.full(500, -405.0, dtype=np.float32)
    mp["energy"] = np.ones(500, dtype=np.float32)
    mp["time"] = np.full(500, 135000.0, dtype=np.float32)
    d = rs.normal(size=(3, 3))
    d[:, 1] = -np.abs(d[:, 1])
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    for j in range(3):
        q = {"x": np.zeros(1), "y": np.full(1, -405.0), "z": np.zeros(1), "dx": d[j:j + 1, 0], "dy": d[j:j + 1, 1],
             "dz": d[j:j + 1, 2], "energy": np.ones(1, dtype=np.float32), "time": np.full(1, 135000.0, np.float32)}
        wf, nbr, rad = oracle.gather(m, abi.BOTTOM_REFLECTION, mp, q, want_nbrs=True)
        mult = (m.floor_albedo / np.pi) * max(0.0, float(np.dot(m.floor_normal, -d[j])))
        expect = mult * m.k / (np.pi * rad[0] ** 2)
        assert wf.sum() == pytest.approx(expect, rel=1e-12)


def test_plane_bounds_the_water_only_below_the_surface():
    """Where the plane rises above the surface (dry ground) a backward photon
    passes through it in the air, records its surface entry and is absorbed:
    no floor record lies above the surface and none follows a dry entry."""
    m = build_scene(tilted_shot(slope=12.0, azimuth=70.0, b=0.5))
    qs, _, _ = oracle.trace(m, (61, 62), abi.PASS_BACKWARD, 0, 400)
    sr = qs[abi.SURFACE_REFLECTION]
    entry = np.stack([sr["x"], sr["y"], sr["z"]], 1).astype(np.float64)
    dry = {int(t >> 16) for t, g in zip(sr["tag"], entry @ m.floor_normal - m.floor_offset) if g <= 0.0}
    assert 20 < len(dry) < 380  # the geometry puts part of the footprint on dry ground
    bt = qs[abi.BOTTOM_REFLECTION]
    assert len(bt["y"]) > 50 and np.all(bt["y"] <= m.y_ws)
    later = {int(t >> 16) for ti in (abi.SCATTER, abi.BOTTOM_REFLECTION) for t in qs[ti]["tag"]}
    assert not (dry & later)
