"""Fixed-radius gather on the GPU vs the oracle's brute force.

For every query the GPU takes exactly the brute-force set (record counts and
sums of record indices equal, query by query) and its FP64 waveform equals
the oracle's sequential re-sum to 1e-9 relative.  Cases:
* footprint grids on the seafloor plane (the BOTTOM map and the SCATTER map),
  in Morton and in index order;
* backward-trace vertices of every photon type;
* a radius that holds the whole map (every subtree taken whole) and a
  one-record map;
* argument errors.
The oracle's radius gather is pinned against scipy's cKDTree and the k-NN
gather in test_oracle_radius.py.
"""
import numpy as np
import pytest

import oracle
from pyalbsim_amd import abi, runtime
from pyalbsim_amd.footprint import footprint_queries, upload_queries
from pyalbsim_amd.model import build_scene
from tests.scenes import layered_scene, run_linear_script_scene, single_shot, tilted_shot

pytestmark = pytest.mark.gpu

F = ("x", "y", "z", "dx", "dy", "dz", "energy", "time")


def _setup(model, n_f, n_b, key):
    ds = runtime.DeviceScene(model)
    recs, rc = runtime.trace(ds, key, abi.PASS_FORWARD, n_f)
    maps = [runtime.build_map(recs[t], rc[t], runtime.MAP_DIMS[t]) for t in range(4)]
    qs, qc = runtime.trace(ds, key, abi.PASS_BACKWARD, n_b)
    return ds, recs, rc, maps, qs, qc


def _host_map(dmap, n):
    return {f: dmap.tensors[f][:n].cpu().numpy() for f in F}


def _check(model, ds, dmap, n, t, sink, qhost, nq, r, sort=True):
    torch = pytest.importorskip("torch")
    wf = torch.zeros(model.n_bins, dtype=torch.float64, device="cuda")
    cnt, isum, _ = runtime.radius_gather(ds, dmap, t, sink, nq, r, wf, model.gate(), counts=True, sort=sort)
    torch.cuda.synchronize()
    ref, rcnt, risum = oracle.radius_gather(model, t, _host_map(dmap, n), {f: qhost[f][:nq] for f in F}, r)
    got_c = cnt.cpu().numpy()[:nq]
    assert np.array_equal(got_c, rcnt), f"type {t}: {np.sum(got_c != rcnt)} of {nq} counts differ"
    assert np.array_equal(isum.cpu().numpy()[:nq], risum), f"type {t}"
    got = wf.cpu().numpy()
    scale = np.max(np.abs(ref))
    if scale == 0:
        assert np.max(np.abs(got)) == 0
    else:
        assert np.max(np.abs(got - ref)) <= 1e-9 * scale, f"type {t}"
    return int(rcnt.sum())


@pytest.mark.parametrize("scene", ["single_shot", "turbid", "layered", "tilted"])
def test_footprint_grid_matches_oracle(gpu, scene):
    cfg = {"single_shot": single_shot, "turbid": run_linear_script_scene, "layered": layered_scene,
           "tilted": lambda: tilted_shot(slope=8.0, azimuth=150.0)}[scene]()
    model = build_scene(cfg)
    ds, recs, rc, maps, qs, qc = _setup(model, 30000, 10, (21, 22))
    fq = footprint_queries(model, grid=24, pitch=0.05)
    sink = upload_queries(fq)
    nq = len(fq["x"])
    taken = _check(model, ds, maps[abi.BOTTOM_REFLECTION], rc[abi.BOTTOM_REFLECTION], abi.BOTTOM_REFLECTION,
                   sink, fq, nq, 0.05)
    assert taken > 1000
    _check(model, ds, maps[abi.BOTTOM_REFLECTION], rc[abi.BOTTOM_REFLECTION], abi.BOTTOM_REFLECTION, sink, fq,
           nq, 0.05, sort=False)
    taken = _check(model, ds, maps[abi.SCATTER], rc[abi.SCATTER], abi.SCATTER, sink, fq, nq, 0.4)
    assert taken > 100


def test_backward_vertices_all_types_match_oracle(gpu):
    model = build_scene(single_shot())
    ds, recs, rc, maps, qs, qc = _setup(model, 20000, 600, (5, 6))
    radius = {abi.BOTTOM_REFLECTION: 0.05, abi.SCATTER: 0.25, abi.SURFACE_REFLECTION: 0.02,
              abi.SURFACE_TRANSMISSION_UP: 0.02}
    for t in range(4):
        nq = min(qc[t], 300)
        if nq == 0 or rc[t] == 0:
            continue
        qhost = {f: qs[t].tensors[f][:nq].cpu().numpy() for f in F}
        _check(model, ds, maps[t], rc[t], t, qs[t], qhost, nq, radius[t])


def test_whole_map_and_single_record(gpu):
    torch = gpu
    model = build_scene(single_shot())
    ds, recs, rc, maps, qs, qc = _setup(model, 5000, 10, (7, 8))
    fq = footprint_queries(model, grid=4, pitch=0.2)
    sink = upload_queries(fq)
    t = abi.SCATTER
    n = rc[t]
    wf = torch.zeros(model.n_bins, dtype=torch.float64, device="cuda")
    cnt, isum, _ = runtime.radius_gather(ds, maps[t], t, sink, 16, 1e4, wf, model.gate(), counts=True)
    assert np.all(cnt.cpu().numpy() == n) and np.all(isum.cpu().numpy() == n * (n - 1) // 2)
    _check(model, ds, maps[t], n, t, sink, fq, 16, 1e4)
    one = runtime.build_map(recs[abi.BOTTOM_REFLECTION], 1, 2)
    _check(model, ds, one, 1, abi.BOTTOM_REFLECTION, sink, fq, 16, 1e4)
    _check(model, ds, one, 1, abi.BOTTOM_REFLECTION, sink, fq, 16, 1e-3)


def test_radius_gather_errors(gpu):
    torch = gpu
    model = build_scene(single_shot())
    ds, recs, rc, maps, qs, qc = _setup(model, 2000, 10, (9, 10))
    sink = upload_queries(footprint_queries(model, grid=2, pitch=0.1))
    wf = torch.zeros(model.n_bins, dtype=torch.float64, device="cuda")
    with pytest.raises(runtime.AlbError, match="radius"):
        runtime.radius_gather(ds, maps[0], 0, sink, 4, 0.0, wf, model.gate())
    with pytest.raises(runtime.AlbError, match="window"):
        runtime.radius_gather(ds, maps[0], 0, sink, 4, 0.1, wf, (0, 4096))
