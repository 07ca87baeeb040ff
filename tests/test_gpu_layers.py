"""Layered and depth-profiled water columns on the GPU vs the CPU oracle.

A layered column is traced one sub-step at a time on both sides, with the
per-depth speed, attenuation, single-scattering albedo and phase function of
physics/models/water.py:253-447 (the oracle's lookups are pinned against
NumPy in test_oracle_layers.py).  Checks:

* trace parity photon by photon (the tolerance of test_gpu_trace_parity:
  positions 1e-5 relative, 99.5 % of photons), forward and backward, with a
  linear or an exponential refractive-index profile, and with absorption,
  scattering and refractive-index profiles of another shape (tabulated);
* a constant layer split into identical layers traces like the closed-form
  single layer (the two GPU paths agree);
* gather parity: neighbour sets bit for bit and waveforms to 1e-9 of the
  oracle's FP64 re-sum, the scatter multiplier taking the phase function of
  the layer at each query's depth;
* the whole path (trace, map, gather) against the oracle on the same streams;
* the batched form with layered and single-layer scenes mixed equals the
  one-scene traces bit for bit and the one-scene waveforms to 1e-9.
"""
import numpy as np
import pytest

import oracle
from pyalbsim_amd import abi, runtime
from pyalbsim_amd.execution.batched import simulate_batched
from pyalbsim_amd.execution.linear import simulate
from pyalbsim_amd.model import build_scene
from pyalbsim_amd.utils.parameter_profile import ExponentialParameter
from tests.scenes import layered_scene, single_shot, stacked_copies
from tests.test_gpu_trace_parity import _compare, _download, _match

pytestmark = pytest.mark.gpu

F = ("x", "y", "z", "dx", "dy", "dz", "energy", "time")


def _layered(kind):
    if kind == "smooth":  # profiles of a, b and n that are neither linear nor exponential (tabulated)
        from tests.scenes import SmoothStepParameter

        return build_scene(layered_scene(n_profile=SmoothStepParameter(1.34, 1.352),
                                         a_profile=SmoothStepParameter(0.04, 0.25),
                                         b_profile=SmoothStepParameter(1.2, 0.3)))
    return build_scene(layered_scene(n_profile=ExponentialParameter(1.34, 1.35) if kind == "exp" else None))


@pytest.mark.parametrize("kind", ["linear", "exp", "smooth"])
@pytest.mark.parametrize("pass_", [abi.PASS_FORWARD, abi.PASS_BACKWARD])
def test_layered_trace_matches_oracle(gpu, kind, pass_):
    counts = _compare(_layered(kind), (91, 92), pass_, 400)
    assert counts[abi.SCATTER] > 100


def test_split_layer_matches_closed_form_legs(gpu):
    cfg = single_shot(a=0.1, b=0.6)
    one = runtime.DeviceScene(build_scene(cfg))
    many = runtime.DeviceScene(build_scene(stacked_copies(cfg, 3)))
    a, ca = runtime.trace(one, (8, 9), abi.PASS_FORWARD, 2000, tags=True)
    b, cb = runtime.trace(many, (8, 9), abi.PASS_FORWARD, 2000, tags=True)
    bad, allp, n = _match(_download(b, cb), _download(a, ca))
    assert n > 2000
    assert 1.0 - len(bad) / len(allp) >= 0.995


def _setup(model, n_f, n_b, key):
    ds = runtime.DeviceScene(model)
    recs, rc = runtime.trace(ds, key, abi.PASS_FORWARD, n_f)
    maps = [runtime.build_map(recs[t], rc[t], runtime.MAP_DIMS[t]) for t in range(4)]
    qs, qc = runtime.trace(ds, key, abi.PASS_BACKWARD, n_b)
    return ds, rc, maps, qs, qc


def test_layered_gather_matches_oracle(gpu):
    torch = gpu
    model = _layered("linear")
    ds, rc, maps, qs, qc = _setup(model, 20000, 1500, (13, 14))
    k = model.k
    win = model.gate()
    for t in range(4):
        if qc[t] == 0 or rc[t] < k:
            continue
        nq = min(qc[t], 600)
        wf = torch.zeros(model.n_bins, dtype=torch.float64, device="cuda")
        nbr = torch.zeros((nq, k), dtype=torch.int32, device="cuda")
        rad = torch.zeros(nq, dtype=torch.float64, device="cuda")
        runtime.gather(ds, maps[t], t, qs[t], nq, wf, win, nbr_out=nbr, radius_out=rad)
        torch.cuda.synchronize()
        mp = {f: maps[t].tensors[f][: rc[t]].cpu().numpy() for f in F}
        qv = {f: qs[t].tensors[f][:nq].cpu().numpy() for f in F}
        if t == abi.SCATTER:
            # queries in every layer, so every layer's phase table is used
            depth = -(qv["y"] - model.y_ws)
            li = np.minimum(np.searchsorted([1.5, 3.5, 6.0], depth, side="right"), 2)
            assert np.all(np.bincount(li, minlength=3) > 10), np.bincount(li)
        qpos = np.stack([qv["x"], qv["y"], qv["z"]], 1)
        idx, _ = oracle.knn(mp["x"], mp["y"], mp["z"], qpos, k)
        assert np.array_equal(nbr.cpu().numpy(), idx), t
        ref = oracle.gather(model, t, mp, qv)
        got = wf.cpu().numpy()
        scale = np.max(np.abs(ref))
        assert scale > 0
        assert np.max(np.abs(got - ref)) <= 1e-9 * scale, f"type {t}"


def test_layered_whole_path_matches_oracle(gpu):
    model = _layered("linear")
    key = (2468, 1357)
    n_f, n_b = 3000, 600
    got = simulate(model, n_f, n_b, key).waveforms.cpu().numpy()
    ref = oracle.run(model, n_f, n_b, key)
    for t in range(4):
        g, r = got[t], ref["waveforms"][t]
        tot = float(np.sum(r))
        if tot == 0:
            assert float(np.sum(g)) == 0
            continue
        assert float(np.sum(g)) == pytest.approx(tot, rel=1e-6), t
        agree = np.abs(g - r) <= 1e-6 * np.maximum(np.abs(r), 1e-300)
        assert np.sum(r[agree]) >= 0.99 * tot, t


def test_batched_mixes_layered_and_single_layer_scenes(gpu):
    models = [_layered("linear"), build_scene(single_shot()), _layered("exp")]
    key = (5150, 5151)
    counts = [2500, 3000, 2000]
    tab = runtime.DeviceSceneTable(models)
    for pass_ in (abi.PASS_FORWARD, abi.PASS_BACKWARD):
        sinks, tot = runtime.trace_batched(tab, key, pass_, counts, tags=True, block=128, refill=2)
        for t in range(4):
            v = sinks[t].view(tot[t])
            sc = v["scene"].cpu().numpy()
            tag = v["tag"].cpu().numpy()
            for s, m in enumerate(models):
                one, oc = runtime.trace(runtime.DeviceScene(m), key, pass_, counts[s], tags=True, scene_id=s)
                ov = one[t].view(oc[t])
                sel = sc == s
                bo, oo = np.argsort(tag[sel]), np.argsort(ov["tag"].cpu().numpy())
                assert np.array_equal(tag[sel][bo], ov["tag"].cpu().numpy()[oo]), (pass_, t, s)
                for f in F:
                    assert np.array_equal(v[f].cpu().numpy()[sel][bo], ov[f].cpu().numpy()[oo]), (pass_, t, s, f)
    nf, nb = [20000, 8000, 12000], [3000, 2000, 2500]
    res = simulate_batched(models, nf, nb, key, win_len=1000)
    for s, m in enumerate(models):
        one = simulate(m, nf[s], nb[s], key, window=m.gate(1000), scene_id=s)
        a, b = res.waveform(s).cpu().numpy(), one.waveforms.cpu().numpy()
        for t in range(4):
            peak = float(np.max(np.abs(b[t])))
            assert np.max(np.abs(a[t] - b[t])) <= 1e-9 * peak + 1e-300, (s, t)


def test_adjoint_rejects_layered_scenes(gpu):
    with pytest.raises(NotImplementedError):
        runtime.DeviceSceneTable([_layered("linear")], adjoint=True)
