"""Adjoint (path replay) against central finite differences with common
random numbers (same Philox key on both sides).

Loss: L = sum over scenes, types and gated bins of R * W with a fixed random
R >= 0, so dL/dW = R exactly.

Exact checks (the perturbed parameter does not move any path, so the CRN
difference is a smooth function of it and must match to rounding):
  * floor albedo (record energies and the bottom factor) and surface
    roughness (GGX BRDF in the gather);
  * absorption a at a fixed attenuation alpha_L (only omega0 changes): pins
    the replay's per-record scatter counts and the record lookup;
  * the HG asymmetry g in the gather's phase-mass table only (sampling table
    held at g).
Statistical check (the parameter moves the paths; both estimators are
unbiased for dE[L]/dtheta): a, b and g, mean over independent keys, within
3 standard errors plus 2 %.
"""
import dataclasses

import numpy as np
import pytest

from pyalbsim_amd import adjoint, runtime
from pyalbsim_amd.execution.batched import simulate_batched
from pyalbsim_amd.sweep import shot_config
from pyalbsim_amd.model import build_scene

pytestmark = pytest.mark.gpu

THETAS = np.array([
    # a,    b,    g,    depth, albedo, roughness
    [0.10, 0.20, 0.90, 5.0, 0.15, 0.30],
    [0.05, 0.60, 0.85, 8.0, 0.30, 0.45],
    [0.25, 1.20, 0.93, 3.0, 0.08, 0.25],
])
NF, NB, WL = 20000, 4000, 1000


def _models(thetas):
    return [adjoint.model_from_theta(t) for t in thetas]


def _weights(models, seed=0):
    rng = np.random.default_rng(seed)
    return [rng.uniform(0.0, 1.0, size=(4, m.gate(WL)[1])) for m in models]


def _loss(res, models, R):
    tot = 0.0
    for s, m in enumerate(models):
        w = res.gated(s, m.gate(WL)).cpu().numpy()
        tot += float(np.sum(w * R[s]))
    return tot


def _dloss(res, models, R):
    torch = pytest.importorskip("torch")
    d = torch.zeros_like(res.waveforms)
    for s, m in enumerate(models):
        g0, gl = m.gate(WL)
        o, nb = int(res.offsets[s]), m.n_bins
        d[o:o + 4 * nb].view(4, nb)[:, g0:g0 + gl] = torch.as_tensor(R[s], device=d.device)
    return d


def _adjoint(models, thetas, key, R):
    res = simulate_batched(models, NF, NB, key, win_len=WL, strict=False, adjoint=True)
    acc, direct = adjoint.adjoint_sums(res, _dloss(res, models, R))
    derivs = [adjoint.scene_derivatives(m, thetas[s]) for s, m in enumerate(models)]
    return adjoint.gradients_from_sums(acc, direct, derivs), acc, direct, _loss(res, models, R)


def _fd(thetas, col, key, R, rel):
    vals = []
    for sign in (1, -1):
        t2 = thetas.copy()
        t2[:, col] *= 1 + sign * rel
        ms = _models(t2)
        # one scene perturbed at a time would cost 3x; the scenes are independent, so perturb all
        r = simulate_batched(ms, NF, NB, key, win_len=WL, strict=False)
        vals.append([float(np.sum(r.gated(s, ms[s].gate(WL)).cpu().numpy() * R[s])) for s in range(len(ms))])
    return (np.array(vals[0]) - np.array(vals[1])) / (2 * rel * thetas[:, col])


@pytest.mark.parametrize("col", [4, 5])
def test_albedo_and_roughness_match_crn_differences(gpu, col):
    models = _models(THETAS)
    R = _weights(models)
    g, _, _, _ = _adjoint(models, THETAS, (7, 8), R)
    fd = _fd(THETAS, col, (7, 8), R, 1e-3)
    assert np.all(np.abs(g[:, col] - fd) <= 2e-3 * np.abs(fd) + 1e-12), (g[:, col], fd)


def test_omega0_weight_matches_crn_difference(gpu):
    """Perturb a with alpha_L held: paths fixed, only omega0 = (alpha - a) / alpha."""
    models = _models(THETAS)
    R = _weights(models, 1)
    _, acc, _, _ = _adjoint(models, THETAS, (9, 10), R)
    pred = np.array([-acc[s, 0] / (m.alpha_l - m.absorption) for s, m in enumerate(models)])
    vals = []
    h = 1e-3
    for sign in (1, -1):
        ms = []
        for m in models:
            a2 = m.absorption * (1 + sign * h)
            ms.append(dataclasses.replace(m, absorption=a2, omega064=(m.alpha_l64 - a2) / m.alpha_l64))
        r = simulate_batched(ms, NF, NB, (9, 10), win_len=WL, strict=False)
        vals.append([float(np.sum(r.gated(s, ms[s].gate(WL)).cpu().numpy() * R[s])) for s in range(len(ms))])
    fd = (np.array(vals[0]) - np.array(vals[1])) / (2 * h * np.array([m.absorption for m in models]))
    assert np.all(np.abs(pred - fd) <= 2e-3 * np.abs(fd) + 1e-12), (pred, fd)


def test_phase_mass_derivative_matches_crn_difference(gpu):
    """Perturb g in the gather's phase-mass table only (sampling unchanged)."""
    models = _models(THETAS)
    R = _weights(models, 2)
    _, _, direct, _ = _adjoint(models, THETAS, (11, 12), R)
    h = 1e-4
    vals = []
    for sign in (1, -1):
        t2 = THETAS.copy()
        t2[:, 2] += sign * h
        gm = _models(t2)
        tab = runtime.DeviceSceneTable(models, gather_phase=gm)
        r = simulate_batched(models, NF, NB, (11, 12), win_len=WL, strict=False, table=tab)
        vals.append([float(np.sum(r.gated(s, models[s].gate(WL)).cpu().numpy() * R[s])) for s in range(3)])
    fd = (np.array(vals[0]) - np.array(vals[1])) / (2 * h)
    assert np.all(np.abs(direct[:, 0] - fd) <= 2e-3 * np.abs(fd) + 1e-12), (direct[:, 0], fd)


def _smooth_weights(models):
    """Slowly varying R: a one-bin shift of a contribution barely changes L,
    which keeps the CRN difference's variance low."""
    out = []
    for m in models:
        gl = m.gate(WL)[1]
        b = np.arange(gl)
        out.append(np.stack([1.0 + 0.5 * np.sin(2 * np.pi * b / 250.0 + t) for t in range(4)]))
    return out


def test_attenuation_and_phase_gradients_agree_with_crn_in_expectation(gpu):
    """a, b, g move the paths: compare means over independent keys (relative
    step 5e-2 for the differences; their curvature error is ~1e-3)."""
    models = _models(THETAS)
    R = _smooth_weights(models)
    G, F = [], []
    for kk in range(16):
        key = (100 + kk, 7)
        g, _, _, _ = _adjoint(models, THETAS, key, R)
        G.append(g[:, :3])
        F.append(np.stack([_fd(THETAS, c, key, R, 5e-2) for c in range(3)], 1))
    G, F = np.array(G), np.array(F)  # [keys, scenes, 3]
    mg, mf = G.mean(0), F.mean(0)
    se = np.sqrt(G.var(0, ddof=1) / len(G) + F.var(0, ddof=1) / len(F))
    ok = np.abs(mg - mf) <= 3 * se + 0.02 * np.abs(mf)
    assert ok.all(), (mg, mf, se)


def test_autograd_function_and_depth(gpu):
    torch = gpu
    th = torch.tensor(THETAS[:2], dtype=torch.float64, requires_grad=True)
    out = adjoint.batched_waveforms(th, NF, NB, (21, 22), win_len=WL, depth_grad="fd")
    assert out.shape == (2, 4, WL)
    models = _models(THETAS[:2])
    R = _weights(models, 4)
    loss = (out * torch.as_tensor(np.array(R), device=out.device)).sum()
    loss.backward()
    g = th.grad.cpu().numpy()
    assert np.all(np.isfinite(g))
    ref, _, _, _ = _adjoint(models, THETAS[:2], (21, 22), R)
    assert np.allclose(g[:, [0, 1, 2, 4, 5]], ref[:, [0, 1, 2, 4, 5]], rtol=1e-9, atol=0)
    fd = _fd(THETAS[:2], 3, (21, 22), R, 1e-3)
    assert np.allclose(g[:, 3], fd, rtol=1e-9)
    # default: one-sided difference from the forward pass's own waveforms
    th1 = torch.tensor(THETAS[:2], dtype=torch.float64, requires_grad=True)
    out1 = adjoint.batched_waveforms(th1, NF, NB, (21, 22), win_len=WL)
    (out1 * torch.as_tensor(np.array(R), device=out1.device)).sum().backward()
    t2 = THETAS[:2].copy()
    t2[:, 3] *= 1 + 1e-3
    ms = _models(t2)
    r = simulate_batched(ms, NF, NB, (21, 22), win_len=WL, strict=False)
    up = np.array([float(np.sum(r.gated(s, ms[s].gate(WL)).cpu().numpy() * R[s])) for s in range(2)])
    base = np.array([float(np.sum(out1[s].detach().cpu().numpy() * R[s])) for s in range(2)])
    assert np.allclose(th1.grad.cpu().numpy()[:, 3], (up - base) / (1e-3 * THETAS[:2, 3]), rtol=1e-9)


def test_independent_run_loss_recipe(gpu):
    """Gradient of the squared error of the mean waveform: the loss is
    evaluated on an independent run (another key) and replayed through the
    first, wf + (wf2 - wf).detach().  Its value is the independent run's loss,
    and since path replay is linear in dL/dW its gradient equals the first
    run's gradient for the fixed weights 2 (wf2 - T)."""
    torch = gpu
    th = torch.tensor(THETAS[:2], dtype=torch.float64, requires_grad=True)
    wf = adjoint.batched_waveforms(th, NF, NB, (21, 22), win_len=WL, depth_grad="none")
    wf2 = adjoint.batched_waveforms(th.detach(), NF, NB, (23, 24), win_len=WL, depth_grad="none")
    assert not torch.equal(wf.detach(), wf2)  # independent photons
    T = 0.5 * wf2.detach() + 0.5 * wf.detach().mean()
    L = ((wf + (wf2 - wf).detach() - T) ** 2).sum()
    lv = float(L.detach())
    assert abs(lv - float(((wf2 - T) ** 2).sum())) <= 1e-10 * abs(lv)
    L.backward()
    th3 = torch.tensor(THETAS[:2], dtype=torch.float64, requires_grad=True)
    wf3 = adjoint.batched_waveforms(th3, NF, NB, (21, 22), win_len=WL, depth_grad="none")
    (wf3 * (2 * (wf2 - T)).detach()).sum().backward()
    g, g3 = th.grad.cpu().numpy(), th3.grad.cpu().numpy()
    assert np.all(np.isfinite(g)) and np.any(g != 0)
    assert np.allclose(g, g3, rtol=1e-9, atol=0)


def _replay_sums(res, weights_by_type, pass_, stat_layer=-1):
    """Run the replay trace of one pass with the given per-slot weights."""
    import ctypes as C

    torch = pytest.importorskip("torch")
    from pyalbsim_amd import abi

    st = res.state
    tab = st["table"]
    L = runtime.load_library()
    nph = st["nf"] if pass_ == abi.PASS_FORWARD else st["nb"]
    sinks = st["records"] if pass_ == abi.PASS_FORWARD else st["queries"]
    counts = st["rcounts"] if pass_ == abi.PASS_FORWARD else st["qcounts"]
    offsets = torch.as_tensor(np.concatenate([[0], np.cumsum(nph)]).astype(np.int64), device="cuda")
    idx = (abi.ReplayIndex * 4)()
    keep = []
    for t in range(4):
        keys, slots, ws = adjoint._replay_index(torch, sinks[t], counts[t], offsets, "cuda")
        w = torch.as_tensor(weights_by_type[t], dtype=torch.float64, device="cuda")
        keep.append((keys, slots, ws, w))
        idx[t].keys, idx[t].slots, idx[t].n, idx[t].weight = runtime._p(keys), runtime._p(slots), counts[t], \
            runtime._p(w)
    acc = torch.zeros(adjoint.REPLAY_SUMS * tab.n, dtype=torch.float64, device="cuda")
    ctr = torch.zeros(abi.N_COUNTERS, dtype=torch.int64, device="cuda")
    work = torch.zeros(1, dtype=torch.int64, device="cuda")
    rng = abi.make_rng(st["key"], pass_)
    rc = L.alb_replay_batched(runtime._p(tab.dev), tab.n, runtime._p(offsets), C.byref(rng), pass_,
                              C.cast(idx, C.c_void_p), runtime._p(work), 128, 2, runtime._p(acc), runtime._p(ctr),
                              int(stat_layer), C.c_void_p(runtime._stream_ptr()))
    assert rc == 0
    torch.cuda.synchronize()
    assert int(ctr[abi.CTR_REPLAY_MISSES].item()) == 0
    return acc.view(tab.n, adjoint.REPLAY_SUMS).cpu().numpy()


@pytest.mark.parametrize("tilt", [None, (9.0, 120.0)])
@pytest.mark.parametrize("pass_", [0, 1])
def test_replay_statistics_match_time_stepped_oracle(gpu, pass_, tilt):
    """The replay trace's per-record path statistics (water length, scatter and
    floor counts, phase-sampling score), weighted by random per-record
    weights, equal the time-stepped CPU oracle's, record by record (matched by
    (photon, event) tag; photons with a one-sub-step rounding flip excluded),
    with the flat floor (closed-form legs) and a tilted one (stepped legs)."""
    import oracle
    from pyalbsim_amd import abi

    thetas = THETAS[:2]
    models = [adjoint.model_from_theta(t, None, tilt) for t in thetas]  # tilt: stepped legs, tilted floor
    nf, nb = [1500, 1200], [600, 500]
    key = (4321, 1234)
    res = simulate_batched(models, nf, nb, key, win_len=WL, strict=False, adjoint=True)
    st = res.state
    sinks = st["records"] if pass_ == abi.PASS_FORWARD else st["queries"]
    counts = st["rcounts"] if pass_ == abi.PASS_FORWARD else st["qcounts"]
    nph = nf if pass_ == abi.PASS_FORWARD else nb
    rng = np.random.default_rng(5)
    gpu_rec = []
    for t in range(4):
        v = sinks[t].view(counts[t])
        gpu_rec.append({"tag": v["tag"].cpu().numpy(), "scene": v["scene"].cpu().numpy(),
                        "y": v["y"].cpu().numpy().astype(np.float64)})
    expect = np.zeros((2, adjoint.REPLAY_SUMS))
    weights = [np.zeros(max(counts[t], 1)) for t in range(4)]
    checked = 0
    for s, m in enumerate(models):
        score, _ = adjoint.phase_g_tables(m.phase_g)
        cpu, _, _ = oracle.trace(m, key, pass_, 0, nph[s], stream=pass_ | (s << 1), stats=True, phase_score=score)
        # photons whose records differ (count or position) between the two traces are left out
        bad = set()
        maps = []
        for t in range(4):
            sel = np.nonzero(gpu_rec[t]["scene"] == s)[0]
            g_by = {int(gpu_rec[t]["tag"][i]): i for i in sel}
            c_by = {int(tg): j for j, tg in enumerate(cpu[t]["tag"])}
            for tg in set(g_by) ^ set(c_by):
                bad.add(tg >> 16)
            for tg in set(g_by) & set(c_by):
                if abs(gpu_rec[t]["y"][g_by[tg]] - float(cpu[t]["y"][c_by[tg]])) > 1e-3:
                    bad.add(tg >> 16)
            maps.append((g_by, c_by))
        for t in range(4):
            g_by, c_by = maps[t]
            for tg, i in g_by.items():
                if (tg >> 16) in bad or tg not in c_by:
                    continue
                j = c_by[tg]
                w = rng.uniform(0.5, 1.5)
                weights[t][i] = w
                c = cpu[t]
                expect[s] += w * np.array([c["nsc"][j], float(c["rdl"][j]) if t == abi.SCATTER else 0.0, c["lam"][j],
                                           c["nfl"][j], c["gs"][j], c["dsc"][j]])
                checked += 1
        assert len(bad) <= 0.01 * nph[s] + 2
    assert checked > 1000
    got = _replay_sums(res, weights, pass_)
    assert np.allclose(got, expect, rtol=1e-9, atol=1e-9), (got, expect)
