"""Whole path on the GPU: drop-in run_linear, agreement with the CPU oracle, and
size-independent invariants at the benchmark's full size.

* Same Philox key on both sides: photon streams are identical, so the
  waveforms agree closely (only FP64 summation order and the rare one-sub-step
  rounding flip differ): the total energy within 1e-6 relative and at least
  99 % of the energy in bins that agree to 1e-6.
* Philox 3 sigma: every gated bin of the GPU run agrees with the oracle run of
  the same key within 3 sigma of the Monte Carlo error (sigma from 32 GPU runs
  with independent keys); an independent-key oracle run agrees statistically.
* 1e8 forward / 2.5e7 backward photons: every photon crosses the surface once
  (one surface record / two surface queries per photon), records lie on their
  planes, energies in (0, 1], event sub-steps inside the budget, waveform
  finite and non-negative.
"""
import numpy as np
import pytest

import oracle
from pyalbsim_amd import abi, runtime
from pyalbsim_amd.execution.linear import simulate
from pyalbsim_amd.model import build_scene
from pyalbsim_amd.photon_mapping.photon_type import PhotonType
from tests.scenes import run_linear_script_scene, single_shot

pytestmark = pytest.mark.gpu


def test_run_linear_dropin_contract(gpu):
    from alb_sim.config.run import RunConfig
    from alb_sim.execution.linear import run_linear

    cfg = run_linear_script_scene()
    rc = RunConfig(photons_per_batch_forward=2000, batches_forward=2, photons_per_batch_backward=500,
                   batches_backward=2)
    wf = run_linear(cfg, rc, np.random.default_rng(7), verbose=False)
    model = build_scene(cfg)
    assert list(wf) == list(PhotonType)
    for v in wf.values():
        assert v.shape == (model.n_bins,) and v.dtype == np.float64
        assert np.all(np.isfinite(v)) and np.all(v >= 0)
    assert sum(float(v.sum()) for v in wf.values()) > 0
    # same generator state -> same key -> same result; batch split does not matter
    rc2 = RunConfig(photons_per_batch_forward=4000, batches_forward=1, photons_per_batch_backward=1000,
                    batches_backward=1)
    wf2 = run_linear(cfg, rc2, np.random.default_rng(7), verbose=False, as_float32=True)
    for t in PhotonType:
        assert wf2[t].dtype == np.float32
        assert np.allclose(wf2[t], wf[t].astype(np.float32), rtol=1e-5, atol=0)


def test_run_linear_replay_mode_matches_oracle(gpu):
    """rng_mode="replay": both passes draw host-supplied uniforms per photon
    slot; the waveforms match the oracle fed the same rows (whole-path
    tolerance), and device_tensors returns the same values on the GPU."""
    from alb_sim.config.run import RunConfig
    from alb_sim.execution.linear import run_linear

    cfg = single_shot()
    rc = RunConfig(photons_per_batch_forward=3000, batches_forward=1, photons_per_batch_backward=600,
                   batches_backward=1)
    rs = np.random.default_rng(99)
    rf, rb = rs.random((3000, 256)), rs.random((600, 256))
    wf = run_linear(cfg, rc, rng_mode="replay", replay_uniforms=(rf, rb), verbose=False)
    ref = oracle.run(build_scene(cfg), 3000, 600, (0, 0), replay_fwd=rf, replay_bwd=rb)
    for pt in PhotonType:
        g, r = wf[pt], ref["waveforms"][pt.abi_index]
        tot = float(np.sum(r))
        if tot == 0:
            assert float(np.sum(g)) == 0
            continue
        assert float(np.sum(g)) == pytest.approx(tot, rel=1e-6), pt
        agree = np.abs(g - r) <= 1e-6 * np.maximum(np.abs(r), 1e-300)
        assert np.sum(r[agree]) >= 0.99 * tot, pt
    dev = run_linear(cfg, rc, rng_mode="replay", replay_uniforms=(rf, rb), device_tensors=True, verbose=False)
    for pt in PhotonType:
        assert dev[pt].is_cuda
        assert np.allclose(dev[pt].cpu().numpy(), wf[pt], rtol=1e-9, atol=1e-12 * max(float(wf[pt].max()), 1e-300))
    with pytest.raises(ValueError):
        run_linear(cfg, rc, rng_mode="replay", verbose=False)


@pytest.mark.parametrize("scene", ["single_shot", "turbid"])
def test_whole_path_matches_oracle_same_streams(gpu, scene):
    model = build_scene(single_shot() if scene == "single_shot" else run_linear_script_scene())
    key = (424242, 17)
    n_f, n_b = (3000, 600) if scene == "single_shot" else (2000, 400)
    res = simulate(model, n_f, n_b, key)
    got = res.waveforms.cpu().numpy()
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


def test_waveform_within_3_sigma_of_independent_runs(gpu):
    """Philox mode: every gated bin of the GPU run agrees with the CPU oracle run
    of the same key within 3 sigma of the Monte Carlo error, sigma taken from
    the spread of 32 GPU runs with independent keys.  An oracle run with an
    independent key is compared statistically: its gated total within 3 sigma
    of the ensemble mean, and >= 90 % of 25-bin groups within 3 sigma (the
    k-NN density estimate is heavy-tailed and correlated across neighbouring
    bins, so a single run can leave 3 sigma in a few adjacent groups)."""
    model = build_scene(single_shot())
    n_f, n_b = 20000, 4000
    g0, glen = model.gate()
    runs = []
    for s in range(32):
        res = simulate(model, n_f, n_b, (1000 + s, 99))
        runs.append(res.waveforms.cpu().numpy()[:, g0:g0 + glen])
    runs = np.array(runs)  # [32, 4, glen]
    sd = runs.std(0, ddof=1)
    # same key: bin by bin
    gpu_same = runs[0]
    cpu_same = np.array([w[g0:g0 + glen] for w in oracle.run(model, n_f, n_b, (1000, 99))["waveforms"]])
    live = sd > 0
    assert np.all(np.abs(gpu_same - cpu_same)[live] <= 3 * sd[live])
    assert np.all(np.abs(gpu_same - cpu_same)[~live] <= 1e-12 * np.abs(cpu_same).max())
    # independent key
    ref = np.array([w[g0:g0 + glen] for w in oracle.run(model, n_f, n_b, (7, 7))["waveforms"]])
    tot = runs.sum(-1)  # [32, 4]
    for t in range(4):
        m, s_ = tot[:, t].mean(), tot[:, t].std(ddof=1)
        # a 3-sigma rule needs a near-Gaussian total: the specular surface return
        # (GGX alpha^2 = roughness^4 = 6.25e-6) is carried by rare near-mirror
        # vertices and varies by orders of magnitude from run to run
        if s_ > 0 and s_ < m:
            assert abs(ref[t].sum() - m) <= 3 * s_ * np.sqrt(1 + 1 / 32), t
    grp = 25
    nb = glen // grp
    R = runs[:, :, :nb * grp].reshape(32, 4, nb, grp).sum(-1)
    C = ref[:, :nb * grp].reshape(4, nb, grp).sum(-1)
    mean = R.mean(0)
    sigma = R.std(0, ddof=1) * np.sqrt(1 + 1 / 32)
    mask = (mean > 1e-6 * mean.max()) & (R.std(0, ddof=1) < R.mean(0))
    z = np.abs(C - mean)[mask] / np.maximum(sigma[mask], 1e-300)
    assert np.mean(z <= 3.0) >= 0.90, np.sort(z)[-5:]


def test_full_size_invariants(gpu):
    """Benchmark size: 1e8 forward and 2.5e7 backward photons."""
    torch = gpu
    model = build_scene(single_shot())
    n_f, n_b = 100_000_000, 25_000_000
    ds = runtime.DeviceScene(model)
    recs, rc = runtime.trace(ds, (5, 5), abi.PASS_FORWARD, n_f)
    assert rc[abi.SURFACE_REFLECTION] == n_f
    f32 = np.float32
    S = model.n_substeps
    for t in range(4):
        v = recs[t].view(rc[t])
        e = v["energy"]
        assert bool(((e > 0) & (e <= 1)).all()), t
        tt = v["time"]
        assert bool(((tt >= 0) & (tt < S)).all()), t
    assert bool((recs[abi.SURFACE_REFLECTION].view(n_f)["y"] == f32(model.y_ws)).all())
    assert bool((recs[abi.SURFACE_TRANSMISSION_UP].view(rc[3])["dy"] > 0).all())
    yb = recs[abi.BOTTOM_REFLECTION].view(rc[0])["y"]
    assert bool((torch.abs(yb.double() - model.y_sf) <= 1e-4).all())
    ys = recs[abi.SCATTER].view(rc[1])["y"]
    assert bool(((ys <= f32(model.y_ws)) & (ys >= f32(model.y_sf))).all())
    del recs
    res = simulate(model, n_f, n_b, (5, 5), keep=True, dscene=ds)
    assert res.query_counts[abi.SURFACE_REFLECTION] == n_b == res.query_counts[abi.SURFACE_TRANSMISSION_UP]
    wf = res.waveforms
    assert bool(torch.isfinite(wf).all()) and bool((wf >= 0).all())
    assert float(wf.sum()) > 0
    assert int(res.counters_fwd[abi.CTR_PATHS]) == n_f and int(res.counters_bwd[abi.CTR_PATHS]) == n_b
    _full_size_neighbour_sets(torch, ds, res)


def _full_size_neighbour_sets(torch, ds, res, n_q=32):
    """On the full-size maps (up to 5e8 records), every type's queries are
    gathered again in one launch that writes each query's neighbour row (up
    to 1.25e8 rows, 5e10 bytes of rows).  n_q random rows per type, with
    their k-th distances, equal brute force over every record bit for bit
    (the index arithmetic of the big launch), and the one-launch waveform
    equals the chunked gather's (launches of 2^24 queries) up to FP64
    summation order, so no chunk drops, repeats or misplaces a query."""
    from concurrent.futures import ThreadPoolExecutor

    k = ds.model.k
    rs = np.random.default_rng(8)
    for t in range(4):
        nq, nm = res.query_counts[t], res.map_counts[t]
        print(f"full-size neighbour rows: type {t}, {nm} records, {nq} queries", flush=True)
        nbr = torch.empty((nq, k), dtype=torch.int32, device="cuda")
        rad = torch.empty(nq, dtype=torch.float64, device="cuda")
        w = torch.zeros(ds.model.n_bins, dtype=torch.float64, device="cuda")
        keep = runtime.gather(ds, res.maps[t], t, res.queries[t], nq, w, ds.model.gate(), nbr_out=nbr,
                              radius_out=rad)
        torch.cuda.synchronize()
        del keep
        ref_w = res.waveforms[t]
        assert float(torch.max(torch.abs(w - ref_w))) <= 1e-9 * float(torch.max(torch.abs(ref_w))), t
        pick = np.sort(rs.choice(nq, size=n_q, replace=False))
        pk = torch.as_tensor(pick, device="cuda")
        got_i = nbr[pk].cpu().numpy()
        got_r = rad[pk].cpu().numpy()
        del nbr, rad
        mx, my, mz = (res.maps[t].tensors[f][:nm].cpu().numpy() for f in ("x", "y", "z"))
        q = np.stack([res.queries[t].tensors[f][pk].cpu().numpy() for f in ("x", "y", "z")], 1)

        def brute(j):
            return oracle.knn(mx, my, mz, q[j:j + 1], k)

        with ThreadPoolExecutor(16) as ex:
            out = list(ex.map(brute, range(n_q)))
        idx = np.concatenate([o[0] for o in out])
        d2 = np.concatenate([o[1] for o in out])
        assert np.array_equal(got_i, idx), t
        assert np.array_equal(got_r, np.sqrt(d2[:, k - 1])), t


def test_run_parallel_contract(gpu):
    """run_parallel returns (waveforms, photon maps, None); the same seed gives
    the same result, and the maps hold the forward records."""
    from alb_sim.config.run import RunConfig
    from alb_sim.execution.parallel import run_parallel

    cfg = single_shot()
    rc = RunConfig(photons_per_batch_forward=3000, batches_forward=2, photons_per_batch_backward=800,
                   batches_backward=2, processes=4)
    wf, maps, heat = run_parallel(cfg, rc, seed=11, verbose=False)
    assert heat is None
    model = build_scene(cfg)
    assert list(wf) == list(PhotonType)
    for pt in PhotonType:
        assert wf[pt].dtype == np.float32 and wf[pt].shape == (model.n_bins,)
        m = maps[pt]
        assert m.positions.shape == (len(m.energies), 3) and m.positions.dtype == np.float32
    assert len(maps[PhotonType.SURFACE_REFLECTION]) == 6000
    assert np.allclose(maps[PhotonType.SURFACE_REFLECTION].positions[:, 1], np.float32(model.y_ws))
    wf2, _, _ = run_parallel(cfg, rc, seed=11, verbose=False)
    for pt in PhotonType:
        assert np.allclose(wf[pt], wf2[pt], rtol=1e-5, atol=0)
