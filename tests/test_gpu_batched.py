"""Batched form (many scenes per launch) against the one-scene path.

Scene s of a batch uses the Philox stream word pass | s << 1, the stream that
runtime.trace(..., scene_id=s) uses for a one-scene trace, so:

* the batched trace's records of scene s equal the one-scene trace's records
  bit for bit (matched by (photon, event) tag), for forward records and
  backward queries, with uneven photon counts and an empty scene;
* the batched map sorts records by (scene, Morton key); every scene range
  [first, last] holds exactly that scene's records and its root node covers
  exactly that range;
* the batched k-NN returns, per query, the same neighbour records (by tag) as
  the one-scene k-NN, and the k-th distance bit for bit;
* the batched waveforms equal the one-scene waveforms up to FP64 summation
  order (relative 1e-9 of the scene's peak bin);
* a scene whose map has fewer than k records raises IndexError (strict) or
  contributes nothing of that type (strict=False), as the reference's
  KDTree.query would fail for it (core/simulation.py:139-142).
The one-scene path itself is checked against the CPU oracle in the other GPU
tests; the oracle consumes the same streams when given the stream word.
"""
import numpy as np
import pytest

import oracle
from pyalbsim_amd import abi, runtime
from pyalbsim_amd.execution.batched import simulate_batched
from pyalbsim_amd.execution.linear import simulate
from pyalbsim_amd.model import build_scene
from pyalbsim_amd.sweep import hypercube
from tests.scenes import single_shot

pytestmark = pytest.mark.gpu

F = ("x", "y", "z", "dx", "dy", "dz", "energy", "time")


def _scenes():
    return [build_scene(single_shot()),
            build_scene(single_shot(depth=12.0, a=0.05, b=0.9, albedo=0.3, roughness=0.3), hg_g=0.9),
            build_scene(single_shot(depth=2.0, a=0.3, b=1.5, albedo=0.05, roughness=0.45), hg_g=0.82)]


def _by_tag(view, n, scene=None, sid=None):
    tag = view["tag"].cpu().numpy()
    sel = np.ones(n, dtype=bool) if scene is None else (scene == sid)
    out = {f: view[f].cpu().numpy()[sel] for f in F}
    order = np.argsort(tag[sel], kind="stable")
    return tag[sel][order], {f: out[f][order] for f in F}


@pytest.mark.parametrize("pass_", [abi.PASS_FORWARD, abi.PASS_BACKWARD])
def test_batched_trace_equals_one_scene_traces(gpu, pass_):
    models = _scenes()
    counts = [3000, 0, 5000, 1777]
    models = [models[0], models[1], models[2], models[1]]
    key = (31337, 4242)
    tab = runtime.DeviceSceneTable(models)
    sinks, tot = runtime.trace_batched(tab, key, pass_, counts, tags=True, block=128, refill=2)
    for t in range(4):
        v = sinks[t].view(tot[t])
        sc = v["scene"].cpu().numpy()
        assert np.all((sc >= 0) & (sc < len(models)))
        for s, m in enumerate(models):
            one, oc = runtime.trace(runtime.DeviceScene(m), key, pass_, counts[s], tags=True, scene_id=s) \
                if counts[s] else (None, [0, 0, 0, 0])
            bt, bv = _by_tag(v, tot[t], sc, s)
            if counts[s] == 0:
                assert len(bt) == 0
                continue
            ot, ov = _by_tag(one[t].view(oc[t]), oc[t])
            assert np.array_equal(bt, ot), (t, s, len(bt), len(ot))
            for f in F:
                assert np.array_equal(bv[f], ov[f]), (t, s, f)


def test_scene_stream_matches_oracle(gpu):
    """The stream word pass | s << 1 is honoured the same way by the oracle."""
    m = _scenes()[1]
    key = (77, 78)
    g, gc = runtime.trace(runtime.DeviceScene(m), key, abi.PASS_FORWARD, 500, tags=True, scene_id=5)
    c, _, _ = oracle.trace(m, key, abi.PASS_FORWARD, 0, 500, stream=abi.PASS_FORWARD | (5 << 1))
    t = abi.SURFACE_REFLECTION
    gt, gv = _by_tag(g[t].view(gc[t]), gc[t])
    co = np.argsort(c[t]["tag"])
    assert np.array_equal(gt, c[t]["tag"][co])
    assert np.allclose(gv["x"], c[t]["x"][co], rtol=1e-5, atol=1e-4)


def test_batched_map_scene_ranges(gpu):
    torch = gpu
    models = _scenes()
    counts = [4000, 2500, 3000]
    tab = runtime.DeviceSceneTable(models)
    sinks, tot = runtime.trace_batched(tab, (5, 6), abi.PASS_FORWARD, counts)
    for t in range(4):
        if tot[t] < 2:
            continue
        m = runtime.build_map_batched(sinks[t], tot[t], runtime.MAP_DIMS[t], len(models))
        torch.cuda.synchronize()
        first = m.tensors["first"].cpu().numpy()
        last = m.tensors["last"].cpu().numpy()
        root = m.tensors["root"].cpu().numpy()
        perm = m.tensors["perm"][:tot[t]].cpu().numpy()
        scene_in = sinks[t].tensors["scene"][:tot[t]].cpu().numpy()
        sorted_scene = scene_in[perm]
        assert np.all(np.diff(sorted_scene) >= 0)
        nodes = m.tensors["nodes"].cpu().numpy()
        for s in range(len(models)):
            cnt = int(np.sum(sorted_scene == s))
            assert last[s] - first[s] + 1 == cnt
            if cnt:
                assert np.all(sorted_scene[first[s]:last[s] + 1] == s)
            if cnt >= 2:
                r = root[s]
                assert nodes[r, 0] == first[s] and nodes[r, 1] == last[s], (t, s)
            elif cnt == 1:
                assert root[s] == ~first[s]


def test_batched_knn_equals_one_scene_knn(gpu):
    torch = gpu
    models = _scenes()
    counts_f = [6000, 3000, 4000]
    counts_b = [600, 500, 700]
    key = (1234, 5678)
    tab = runtime.DeviceSceneTable(models)
    recs, rt = runtime.trace_batched(tab, key, abi.PASS_FORWARD, counts_f, tags=True)
    qs, qt = runtime.trace_batched(tab, key, abi.PASS_BACKWARD, counts_b, tags=True)
    k = tab.k
    for t in (abi.SCATTER, abi.BOTTOM_REFLECTION):
        bm = runtime.build_map_batched(recs[t], rt[t], runtime.MAP_DIMS[t], len(models))
        nb = torch.empty((qt[t], k), dtype=torch.int32, device="cuda")
        rb = torch.empty(qt[t], dtype=torch.float64, device="cuda")
        wf = torch.zeros(sum(4 * m.n_bins for m in models), dtype=torch.float64, device="cuda")
        offs = torch.as_tensor(np.concatenate([[0], np.cumsum([4 * m.n_bins for m in models])[:-1]]).astype(np.int64),
                               device="cuda")
        win0 = torch.zeros(len(models), dtype=torch.int32, device="cuda")
        keep = runtime.gather_batched(tab, bm, t, qs[t], qt[t], wf, offs, win0, 0, nbr_out=nb, radius_out=rb)
        order = keep[1].cpu().numpy()
        nb_h, rb_h = nb.cpu().numpy(), rb.cpu().numpy()
        qtag = qs[t].tensors["tag"][:qt[t]].cpu().numpy()
        qsc = qs[t].tensors["scene"][:qt[t]].cpu().numpy()
        mtag = recs[t].tensors["tag"][:rt[t]].cpu().numpy()[bm.tensors["perm"][:rt[t]].cpu().numpy()]
        for s, m in enumerate(models):
            ds = runtime.DeviceScene(m)
            r1, c1 = runtime.trace(ds, key, abi.PASS_FORWARD, counts_f[s], tags=True, scene_id=s)
            q1, d1 = runtime.trace(ds, key, abi.PASS_BACKWARD, counts_b[s], tags=True, scene_id=s)
            if c1[t] < k or d1[t] == 0:
                continue
            pm = runtime.build_map(r1[t], c1[t], runtime.MAP_DIMS[t])
            n1 = torch.empty((d1[t], k), dtype=torch.int32, device="cuda")
            rr1 = torch.empty(d1[t], dtype=torch.float64, device="cuda")
            w1 = torch.zeros(m.n_bins, dtype=torch.float64, device="cuda")
            runtime.gather(ds, pm, t, q1[t], d1[t], w1, (0, 0), nbr_out=n1, radius_out=rr1)
            torch.cuda.synchronize()
            ptag = r1[t].tensors["tag"][:c1[t]].cpu().numpy()[pm.tensors["perm"][:c1[t]].cpu().numpy()]
            q1tag = q1[t].tensors["tag"][:d1[t]].cpu().numpy()
            one = {int(q1tag[i]): (np.sort(ptag[n1.cpu().numpy()[i]]), rr1.cpu().numpy()[i]) for i in range(d1[t])}
            slots = np.nonzero(qsc[order] == s)[0]
            assert len(slots) == d1[t]
            for si in slots:
                qi = order[si]
                got = np.sort(mtag[nb_h[si]])
                ref, rad = one[int(qtag[qi])]
                assert np.array_equal(got, ref), (t, s, si)
                assert rb_h[si] == rad


def test_simulate_batched_equals_one_scene(gpu):
    models = _scenes()
    nf = [20000, 8000, 12000]
    nb = [3000, 2000, 2500]
    key = (99, 100)
    res = simulate_batched(models, nf, nb, key, win_len=1000)
    for s, m in enumerate(models):
        one = simulate(m, nf[s], nb[s], key, window=m.gate(1000), scene_id=s)
        a = res.waveform(s).cpu().numpy()
        b = one.waveforms.cpu().numpy()
        assert a.shape == b.shape
        for t in range(4):
            peak = float(np.max(np.abs(b[t])))
            assert np.max(np.abs(a[t] - b[t])) <= 1e-9 * peak + 1e-300, (s, t)
        assert list(res.scene_map_counts[s]) == list(one.map_counts)


def test_small_scene_error_semantics(gpu):
    models = _scenes()[:2]
    nf, nb = [20000, 30], [2000, 300]
    with pytest.raises((IndexError, KeyError)):
        simulate_batched(models, nf, nb, (1, 2))
    res = simulate_batched(models, nf, nb, (1, 2), strict=False)
    one = simulate(models[0], nf[0], nb[0], (1, 2), window=models[0].gate(1024), scene_id=0)
    a, b = res.waveform(0).cpu().numpy(), one.waveforms.cpu().numpy()
    assert np.allclose(a, b, rtol=1e-9, atol=1e-12 * float(np.abs(b).max()))


def test_hypercube_batch_runs(gpu):
    """64 sweep scenes at 1/1000 of their photon counts: finite, non-negative
    waveforms and one surface record / two surface queries per photon."""
    sc = hypercube(64, seed=1, log10_photons=(2.0, 4.0))
    res = simulate_batched([x.model for x in sc], [x.n_forward for x in sc], [x.n_backward for x in sc], (3, 1),
                           strict=False)
    torch = gpu
    assert bool(torch.isfinite(res.waveforms).all()) and bool((res.waveforms >= 0).all())
    assert res.map_counts[abi.SURFACE_REFLECTION] == sum(x.n_forward for x in sc)
    assert res.query_counts[abi.SURFACE_REFLECTION] == sum(x.n_backward for x in sc)
    assert int(res.counters_fwd[abi.CTR_PATHS]) == sum(x.n_forward for x in sc)


def test_autotuned_launch_gives_same_waveforms(gpu):
    """Block size and refill interval change speed only: the autotuned run
    equals a fixed-parameter run up to FP64 summation order."""
    from pyalbsim_amd.autotune import tune_batched

    models = _scenes()
    nf, nb = [8000, 5000, 6000], [1500, 1000, 1200]
    b, r, times = tune_batched(models, nf, (3, 4))
    assert b in (64, 128, 256) and r in (1, 2, 4, 8, 16) and len(times) == 15
    a = simulate_batched(models, nf, nb, (3, 4), block="auto")
    c = simulate_batched(models, nf, nb, (3, 4), block=64, refill=1)
    for s in range(3):
        x, y = a.waveform(s).cpu().numpy(), c.waveform(s).cpu().numpy()
        assert np.allclose(x, y, rtol=1e-9, atol=1e-12 * float(np.abs(y).max()))


def test_grouped_sweep_equals_one_group(gpu):
    """run_sweep packs scenes into contiguous memory-bounded groups; scene s
    keeps stream id s, so any grouping gives the one-group waveforms up to
    FP64 summation order."""
    from pyalbsim_amd.sweep import estimate_group_bytes, run_sweep

    sc = hypercube(24, seed=1, log10_photons=(2.5, 3.5))
    key = (9, 2)
    models = [x.model for x in sc]
    nf, nb = [x.n_forward for x in sc], [x.n_backward for x in sc]
    est = estimate_group_bytes(models, nf, nb, key)
    assert np.all(est > 0)
    budget = float(est.sum()) / 5.0
    res = run_sweep(sc, key, budget_bytes=budget, margin=1.0, strict=False)
    assert len(res.groups) >= 4 and res.groups[0][0] == 0 and res.groups[-1][1] == len(sc)
    assert all(a < b and b == c for (a, b), (c, _) in zip(res.groups, res.groups[1:]))
    one = simulate_batched(models, nf, nb, key, win_len=1000, strict=False)
    for s in range(len(sc)):
        ref = one.gated(s, res.windows[s]).cpu().numpy()
        got = res.waveforms[s]
        assert got.shape == ref.shape
        assert np.allclose(got, ref, rtol=1e-9, atol=1e-12 * max(float(np.abs(ref).max()), 1e-300)), s
