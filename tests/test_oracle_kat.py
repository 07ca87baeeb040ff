"""Known-answer tests of the oracle's stepping, handlers and gather.

* Ballistic scene: particle scattering switched off (b equal to the pure-water
  value, a = 1e-12) makes every path straight until the floor, so the surface
  and floor records follow from ray-plane geometry, Snell's law and the
  Schlick transmission alone.
* k-NN: brute force against scipy.spatial.cKDTree (an independent k-d tree).
* Gather: FP64 re-sum with scipy neighbours, and the density-estimate
  invariant k / (pi r_k^2) ~ rho (planar) and k / (4/3 pi r_k^3) ~ rho
  (volume) on regular lattices, which fixes the kernel normalisation.
"""
import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle
from pyalbsim_amd import abi
from pyalbsim_amd.model import build_scene
from tests.scenes import single_shot

BW = (1 + 0.008027 * 37.0) * 0.00012 * 0.532 ** (-4.24)


def _snell(d, eta):
    d = d / np.linalg.norm(d)
    n = np.array([0.0, 1.0, 0.0])
    ci = -d @ n
    ct = math.sqrt(1 - eta ** 2 * (1 - ci ** 2))
    t = eta * d + (eta * ci - ct) * n
    return t / np.linalg.norm(t)


@pytest.mark.parametrize("nadir", [0.0, 12.0])
def test_ballistic_forward_records(nadir):
    m = build_scene(single_shot(a=1e-12, b=BW, nadir=nadir))
    assert m.alpha_l == pytest.approx(1e-12, rel=1e-6)
    n = 150
    key = (31, 41)
    recs, ctr, ex = oracle.trace(m, key, abi.PASS_FORWARD, 0, n)
    assert ex == 0
    assert len(recs[abi.SCATTER]["x"]) == 0
    sr = recs[abi.SURFACE_REFLECTION]
    assert len(sr["x"]) == n
    d0 = oracle.emission_dirs(m, key, abi.PASS_FORWARD, n)
    photon = sr["tag"] >> 16
    for j in range(n):
        d = d0[photon[j]]
        P = d * (m.y_ws / d[1])
        assert np.allclose([sr["x"][j], sr["y"][j], sr["z"][j]], P, atol=1e-4)
        assert sr["energy"][j] == 1.0
        dy = (d[1] * m.v_air) * m.dt
        s_pred = math.ceil(m.y_ws / dy) - 1
        assert abs(sr["time"][j] - s_pred) <= 1
    # first floor hit of every photon: refracted straight ray, energy T
    bt = recs[abi.BOTTOM_REFLECTION]
    first = (bt["tag"] & 0xFFFF) == 1
    assert first.sum() == n
    for j in np.nonzero(first)[0]:
        p = bt["tag"][j] >> 16
        d = d0[p]
        P = d * (m.y_ws / d[1])
        t = _snell(d, 1 / m.n_water)
        Q = P + t * ((m.y_sf - P[1]) / t[1])
        assert np.allclose([bt["x"][j], bt["y"][j], bt["z"][j]], Q, atol=1e-4)
        ci = -d[1]
        T = 1 - (m.F0 + (1 - m.F0) * (1 - ci) ** 5)
        assert bt["energy"][j] == np.float32(T)
        s_enter = math.ceil(m.y_ws / ((d[1] * m.v_air) * m.dt)) - 1
        water_steps = (m.y_sf - m.y_ws) / ((t[1] * m.v_water) * m.dt)
        assert abs(bt["time"][j] - (s_enter + water_steps)) <= 2
    # transmitted-up records carry energy T * albedo (float32) and lie on the surface
    tu = recs[abi.SURFACE_TRANSMISSION_UP]
    assert 0 < len(tu["x"]) <= len(bt["x"])
    assert np.all(tu["y"] == np.float32(m.y_ws))
    assert np.all(tu["dy"] > 0)


def test_ballistic_backward_queries():
    m = build_scene(single_shot(a=1e-12, b=BW))
    n = 100
    key = (5, 8)
    qs, ctr, ex = oracle.trace(m, key, abi.PASS_BACKWARD, 0, n)
    d0 = oracle.emission_dirs(m, key, abi.PASS_BACKWARD, n)
    for t in (abi.SURFACE_REFLECTION, abi.SURFACE_TRANSMISSION_UP):
        q = qs[t]
        assert len(q["x"]) == n
        for j in range(n):
            d = d0[q["tag"][j] >> 16]
            P = d * (m.y_ws / d[1])
            assert np.allclose([q["x"][j], q["y"][j], q["z"][j]], P, atol=1e-8)
            assert np.allclose([q["dx"][j], q["dy"][j], q["dz"][j]], d, atol=0)


def test_knn_matches_scipy_kdtree():
    rs = np.random.default_rng(8)
    pts = rs.normal(size=(3000, 3)).astype(np.float32)
    pts[:1000, 1] = np.float32(-400.0)  # a planar subset like the surface maps
    q = rs.normal(size=(200, 3)) * 1.5
    idx, d2 = oracle.knn(pts[:, 0], pts[:, 1], pts[:, 2], q, 100)
    tree = cKDTree(pts.astype(np.float64))
    dd, ii = tree.query(q, k=100)
    assert np.array_equal(np.sort(idx, 1), np.sort(ii, 1))
    assert np.allclose(np.sqrt(d2), dd, rtol=1e-12)


def test_knn_orders_ties_by_index():
    pts = np.zeros((10, 3), dtype=np.float32)
    pts[:, 0] = [1, -1, 1, -1, 2, 2, 3, 0.5, -0.5, 0.5]
    idx, d2 = oracle.knn(pts[:, 0], pts[:, 1], pts[:, 2], np.zeros((1, 3)), 5)
    assert idx[0].tolist() == [7, 8, 9, 0, 1]
    with pytest.raises(IndexError):
        oracle.knn(pts[:, 0], pts[:, 1], pts[:, 2], np.zeros((1, 3)), 11)


def _map(x, y, z, d, e=None, t=None):
    n = len(x)
    return {"x": np.asarray(x, np.float32), "y": np.asarray(y, np.float32), "z": np.asarray(z, np.float32),
            "dx": np.full(n, d[0], np.float32), "dy": np.full(n, d[1], np.float32),
            "dz": np.full(n, d[2], np.float32),
            "energy": np.ones(n, np.float32) if e is None else np.asarray(e, np.float32),
            "time": np.zeros(n, np.float32) if t is None else np.asarray(t, np.float32)}


def _queries(p, d, e=1.0, t=0.0):
    p = np.atleast_2d(p)
    n = len(p)
    return {"x": p[:, 0], "y": p[:, 1], "z": p[:, 2], "dx": np.full(n, d[0]), "dy": np.full(n, d[1]),
            "dz": np.full(n, d[2]), "energy": np.full(n, e, np.float32), "time": np.full(n, t, np.float32)}


def test_planar_kernel_density_invariant():
    """BOTTOM gather on a square lattice of spacing h: sum over the k
    neighbours of albedo/pi / (pi r_k^2) ~ rho * albedo/pi with rho = 1/h^2."""
    m = build_scene(single_shot())
    h = 0.01
    g = np.arange(-60, 61) * h
    X, Z = np.meshgrid(g, g)
    mp = _map(X.ravel(), np.full(X.size, m.y_sf), Z.ravel(), (0.0, -1.0, 0.0))
    q = _queries([[0.0013, m.y_sf, -0.0021]], (0.0, -1.0, 0.0))
    wf = oracle.gather(m, abi.BOTTOM_REFLECTION, mp, q, n_bins=16)
    est = wf[0] / (m.floor_albedo / np.pi)
    assert est == pytest.approx(1 / h ** 2, rel=0.12)
    assert np.count_nonzero(wf) == 1


def test_volume_kernel_density_invariant():
    m = build_scene(single_shot())
    h = 0.02
    g = np.arange(-12, 13) * h
    X, Y, Z = np.meshgrid(g, g + m.y_ws - 1.0, g)
    mp = _map(X.ravel(), Y.ravel(), Z.ravel(), (0.0, 0.0, 1.0))
    q = _queries([[0.001, m.y_ws - 1.0 + 0.002, -0.003]], (0.0, 0.0, -1.0))
    wf = oracle.gather(m, abi.SCATTER, mp, q, n_bins=16)
    m0 = oracle.scatter_mult(np.array([0, 0, 1], np.float32), np.array([0.0, 0.0, -1.0]), m.phase_theta, m.phase_cdf)
    assert m0 > 0
    est = wf[0] / m0
    assert est == pytest.approx(1 / h ** 3, rel=0.15)


def test_gather_equals_fp64_resum_with_scipy_neighbours():
    m = build_scene(single_shot())
    rs = np.random.default_rng(9)
    n = 2000
    dirs = rs.normal(size=(n, 3))
    dirs[:, 1] = np.abs(dirs[:, 1])
    mp = {"x": rs.normal(size=n).astype(np.float32), "y": np.full(n, m.y_ws, np.float32),
          "z": rs.normal(size=n).astype(np.float32), "dx": dirs[:, 0].astype(np.float32),
          "dy": dirs[:, 1].astype(np.float32), "dz": dirs[:, 2].astype(np.float32),
          "energy": rs.random(n).astype(np.float32),
          "time": rs.integers(133000, 134000, n).astype(np.float32)}
    nq = 50
    qd = np.column_stack([rs.normal(size=nq) * 0.01, -np.ones(nq), rs.normal(size=nq) * 0.01])
    qs = {"x": rs.normal(size=nq), "y": np.full(nq, m.y_ws), "z": rs.normal(size=nq), "dx": qd[:, 0],
          "dy": qd[:, 1], "dz": qd[:, 2], "energy": rs.random(nq).astype(np.float32),
          "time": rs.integers(133000, 134000, nq).astype(np.float32)}
    tree = cKDTree(np.column_stack([mp["x"], mp["y"], mp["z"]]).astype(np.float64))
    for t in (abi.SURFACE_REFLECTION, abi.SURFACE_TRANSMISSION_UP):
        got = oracle.gather(m, t, mp, qs)
        ref = np.zeros(m.n_bins)
        for j in range(nq):
            dist, idx = tree.query([qs["x"][j], qs["y"][j], qs["z"][j]], k=m.k)
            knorm = 1.0 / (np.pi * dist[-1] ** 2)
            ds = np.array([qs["dx"][j], qs["dy"][j], qs["dz"][j]])
            for i in idx:
                dp = np.array([mp["dx"][i], mp["dy"][i], mp["dz"][i]], np.float32)
                if t == abi.SURFACE_REFLECTION:
                    mul = oracle.brdf_reflect(dp, ds, m.roughness ** 2, m.F0)
                else:
                    mul = oracle.btdf_transmit(dp, ds, m.roughness ** 2, m.n_water, m.F0, m.surface_albedo)
                e = float(np.float32(qs["energy"][j] * mp["energy"][i])) * knorm * mul
                b = int(np.float32(qs["time"][j] + mp["time"][i]) / np.float32(m.mult))
                ref[b] += e
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref))
