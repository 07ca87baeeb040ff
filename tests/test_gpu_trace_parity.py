"""HIP trace vs the time-stepped CPU oracle, photon by photon.

Both sides consume identical per-photon uniform streams (Philox4x32-10 or a
replayed host buffer).  The oracle advances every photon one sub-step at a
time like src/alb_sim/core/simulation.py; the HIP kernel solves legs in
closed form.  Records are matched by (photon, event) tag.

Tolerance (trajectories): |dpos| <= 1e-5 |pos| (positions are ~400 m from the
origin, FP32 records), direction components within 1e-5, energies within
1e-5 relative, event sub-step within 1.  At least 99.5 % of the photons must
match on every record; the rest may differ by a one-sub-step rounding flip of
a crossing (a 1e-11-relative position difference decides it).
"""
import numpy as np
import pytest

import oracle
from pyalbsim_amd import abi, runtime
from pyalbsim_amd.model import build_scene
from tests.scenes import run_linear_script_scene, single_shot

pytestmark = pytest.mark.gpu

FIELDS = ("x", "y", "z", "dx", "dy", "dz", "energy", "time", "tag")


def _download(sinks, counts):
    out = []
    for t in range(4):
        v = sinks[t].view(counts[t])
        out.append({f: v[f].cpu().numpy() for f in FIELDS})
    return out


def _match(gpu_sets, cpu_sets):
    """Per type: fraction of photons whose records all agree, and details."""
    bad_photons = set()
    all_photons = set()
    n_checked = 0
    for t in range(4):
        g, c = gpu_sets[t], cpu_sets[t]
        all_photons.update((c["tag"] >> 16).tolist())
        all_photons.update((g["tag"] >> 16).tolist())
        gi = {int(tag): i for i, tag in enumerate(g["tag"])}
        ci = {int(tag): i for i, tag in enumerate(c["tag"])}
        for tag in set(gi) ^ set(ci):
            bad_photons.add(tag >> 16)
        for tag in set(gi) & set(ci):
            a, b = gi[tag], ci[tag]
            pg = np.array([g["x"][a], g["y"][a], g["z"][a]], dtype=np.float64)
            pc = np.array([c["x"][b], c["y"][b], c["z"][b]], dtype=np.float64)
            dg = np.array([g["dx"][a], g["dy"][a], g["dz"][a]], dtype=np.float64)
            dc = np.array([c["dx"][b], c["dy"][b], c["dz"][b]], dtype=np.float64)
            ok = np.linalg.norm(pg - pc) <= 1e-5 * max(np.linalg.norm(pc), 1.0)
            ok &= np.max(np.abs(dg - dc)) <= 1e-5
            ok &= abs(float(g["energy"][a]) - float(c["energy"][b])) <= 1e-5 * abs(float(c["energy"][b])) + 1e-30
            ok &= abs(float(g["time"][a]) - float(c["time"][b])) <= 1.0
            n_checked += 1
            if not ok:
                bad_photons.add(tag >> 16)
    return bad_photons, all_photons, n_checked


def _compare(model, key, pass_, n, replay=None, min_frac=0.995):
    import torch

    ds = runtime.DeviceScene(model)
    rep_dev = torch.as_tensor(replay, device="cuda") if replay is not None else None
    sinks, counts = runtime.trace(ds, key, pass_, n, replay=rep_dev, tags=True)
    g = _download(sinks, counts)
    c, ctr, ex = oracle.trace(model, key, pass_, 0, n, replay=replay)
    assert ex == 0
    for t in range(4):
        # identical record counts per type up to the rare rounding flips
        assert abs(len(c[t]["x"]) - counts[t]) <= max(2, 0.005 * len(c[t]["x"])), (t, len(c[t]["x"]), counts[t])
    bad, allp, n_checked = _match(g, c)
    assert n_checked > 0
    frac_ok = 1.0 - len(bad) / max(len(allp), 1)
    assert frac_ok >= min_frac, f"{len(bad)} of {len(allp)} photons disagree: {sorted(bad)[:10]}"
    return counts


@pytest.mark.parametrize("pass_", [abi.PASS_FORWARD, abi.PASS_BACKWARD])
def test_trace_philox_matches_oracle(gpu, pass_):
    model = build_scene(single_shot())
    counts = _compare(model, (12345, 678), pass_, 400)
    assert counts[abi.SURFACE_REFLECTION if pass_ == abi.PASS_FORWARD else abi.SURFACE_TRANSMISSION_UP] > 0


def test_trace_turbid_scene_matches_oracle(gpu):
    """scripts/run_linear.py scene: 15 degree nadir, 7 m, b = 3.5/m (many scatters)."""
    model = build_scene(run_linear_script_scene())
    _compare(model, (7, 9), abi.PASS_FORWARD, 300)
    _compare(model, (7, 9), abi.PASS_BACKWARD, 200)


@pytest.mark.parametrize("pass_", [abi.PASS_FORWARD, abi.PASS_BACKWARD])
def test_trace_replay_matches_oracle(gpu, pass_):
    """Replay mode: host-supplied FP64 uniforms, 256 per photon slot."""
    model = build_scene(single_shot())
    rs = np.random.default_rng(2024 + pass_)
    replay = rs.random((300, 256))
    _compare(model, (0, 0), pass_, 300, replay=replay)


def test_trace_is_deterministic(gpu):
    model = build_scene(single_shot())
    ds = runtime.DeviceScene(model)
    a, ca = runtime.trace(ds, (5, 6), abi.PASS_FORWARD, 20000, tags=True)
    b, cb = runtime.trace(ds, (5, 6), abi.PASS_FORWARD, 20000, tags=True)
    assert ca == cb
    for t in range(4):
        ta = a[t].view(ca[t])
        tb = b[t].view(cb[t])
        oa = np.argsort(ta["tag"].cpu().numpy())
        ob = np.argsort(tb["tag"].cpu().numpy())
        for f in ("x", "y", "z", "energy", "time"):
            assert np.array_equal(ta[f].cpu().numpy()[oa], tb[f].cpu().numpy()[ob])
