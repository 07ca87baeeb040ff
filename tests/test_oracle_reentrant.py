"""The oracle keeps no mutable state between calls: traces and gathers run
from several threads at once give what the same calls give one after another,
heat maps and interaction points included.  The gather over neighbour sets
from a k-d tree equals the brute-force gather."""
from concurrent.futures import ThreadPoolExecutor

import numpy as np
from scipy.spatial import cKDTree

import oracle
from pyalbsim_amd import abi
from pyalbsim_amd.model import build_scene
from tests.scenes import run_linear_script_scene, single_shot

HEAT = {"bins": 40, "water_extent": 2.0, "seafloor_extent": 4.0}


def _trace(args):
    model, begin, n = args
    h = dict(HEAT)
    sinks, ctr, _ = oracle.trace(model, (3, 9), abi.PASS_FORWARD, begin, n, points=True, heat=h)
    return sinks, ctr, h["radius"]


def test_concurrent_traces_equal_sequential():
    model = build_scene(run_linear_script_scene())
    jobs = [(model, b, 150) for b in range(0, 1200, 150)]
    seq = [_trace(j) for j in jobs]
    with ThreadPoolExecutor(8) as ex:
        par = list(ex.map(_trace, jobs))
    for (s1, c1, r1), (s2, c2, r2) in zip(seq, par):
        assert np.array_equal(c1, c2) and np.array_equal(r1, r2)
        for t in range(4):
            for f, v in s1[t].items():
                assert np.array_equal(v, s2[t][f], equal_nan=True), (t, f)


def test_concurrent_gathers_with_heat_maps_equal_sequential():
    model = build_scene(single_shot())
    maps, _, _ = oracle.trace(model, (4, 4), abi.PASS_FORWARD, 0, 3000, points=True)
    qs, _, _ = oracle.trace(model, (4, 4), abi.PASS_BACKWARD, 0, 300)
    t = abi.BOTTOM_REFLECTION
    chunks = [{f: v[b:b + 40] for f, v in qs[t].items()} for b in range(0, len(qs[t]["x"]), 40)]

    def run(q):
        h = dict(HEAT)
        wf = oracle.gather(model, t, maps[t], q, heat=h)
        return wf, h["water"], h["floor"]

    seq = [run(q) for q in chunks]
    with ThreadPoolExecutor(8) as ex:
        par = list(ex.map(run, chunks))
    for a, b in zip(seq, par):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)


def test_gather_over_given_neighbour_sets_equals_brute_force():
    model = build_scene(single_shot())
    maps, _, _ = oracle.trace(model, (5, 5), abi.PASS_FORWARD, 0, 4000)
    qs, _, _ = oracle.trace(model, (5, 5), abi.PASS_BACKWARD, 0, 400)
    for t in range(4):
        mp, q = maps[t], qs[t]
        if len(q["x"]) == 0:
            continue
        pts = np.stack([mp["x"], mp["y"], mp["z"]], 1).astype(np.float64)
        qp = np.stack([q["x"], q["y"], q["z"]], 1)
        _, nb = cKDTree(pts).query(qp, k=model.k)
        ref, nb_bf, rad = oracle.gather(model, t, mp, q, want_nbrs=True)
        # the tree's sets are the brute-force sets here (no ties at the k-th distance)
        assert np.array_equal(np.sort(nb, axis=1), np.sort(nb_bf, axis=1))
        got = oracle.gather(model, t, mp, q, nbrs=nb)
        assert np.max(np.abs(got - ref)) <= 1e-12 * np.max(np.abs(ref)), t
