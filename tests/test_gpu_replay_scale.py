"""Replay mode at the 1e8 workload's replay sizes.

Both readings of "1e5 host-supplied uniform draws per photon slot":
  (i)  1e5 photon slots x 256 FP64 uniforms: the forward trace of all slots
       on the GPU; 400 sampled slots re-traced by the oracle from the same
       rows must match record by record (trajectory tolerance of
       test_gpu_trace_parity: positions 1e-5 relative, directions 1e-5,
       energies 1e-5 relative, sub-step +-1, 99.5 % of photons); no slot
       runs out of uniforms.  The maps of these records and the backward
       queries of 2.5e4 replayed sensor slots then give k-NN neighbour sets
       equal to the brute force bit for bit (300 sampled queries per type).
  (ii) 1e3 photon slots x 1e5 uniforms: every photon against the oracle.
"""
import numpy as np
import pytest

import oracle
from pyalbsim_amd import abi, runtime
from pyalbsim_amd.model import build_scene
from tests.scenes import single_shot
from tests.test_gpu_trace_parity import _compare, _match

pytestmark = pytest.mark.gpu

F = ("x", "y", "z", "dx", "dy", "dz", "energy", "time", "tag")


def _select(view, n, photons):
    tag = view["tag"][:n].cpu().numpy()
    sel = np.isin(tag >> 16, photons)
    return {f: view[f][:n].cpu().numpy()[sel] for f in F}


def test_replay_1e5_slots_256_uniforms(gpu):
    torch = gpu
    model = build_scene(single_shot())
    rs = np.random.default_rng(777)
    n = 100_000
    rep = rs.random((n, 256))
    ds = runtime.DeviceScene(model)
    ctr = torch.zeros(abi.N_COUNTERS, dtype=torch.int64, device="cuda")
    rep_dev = torch.as_tensor(rep, device="cuda")
    sinks, counts = runtime.trace(ds, (0, 0), abi.PASS_FORWARD, n, replay=rep_dev, tags=True, counters=ctr)
    assert int(ctr[abi.CTR_REPLAY_EXHAUSTED].item()) == 0
    assert int(ctr[abi.CTR_PATHS].item()) == n
    sample = np.sort(rs.choice(n, 400, replace=False))
    gpu_sets = [_select(sinks[t].tensors, counts[t], sample) for t in range(4)]
    cpu_sets = [{f: [] for f in F} for _ in range(4)]
    for j in sample:
        recs, _, ex = oracle.trace(model, (0, 0), abi.PASS_FORWARD, int(j), 1, replay=rep[j:j + 1])
        assert ex == 0
        for t in range(4):
            for f in F:
                cpu_sets[t][f].append(recs[t][f])
    cpu_sets = [{f: np.concatenate(c[f]) for f in F} for c in cpu_sets]
    bad, allp, n_checked = _match(gpu_sets, cpu_sets)
    assert n_checked >= 400
    assert 1.0 - len(bad) / len(allp) >= 0.995, sorted(bad)[:10]

    # neighbour sets over the replayed maps, replayed sensor paths
    maps = [runtime.build_map(sinks[t], counts[t], runtime.MAP_DIMS[t]) for t in range(4)]
    nb = 25_000
    repb = torch.as_tensor(rs.random((nb, 256)), device="cuda")
    qs, qc = runtime.trace(ds, (0, 0), abi.PASS_BACKWARD, nb, replay=repb)
    k = model.k
    checked = 0
    for t in range(4):
        nq = min(qc[t], 300)
        if nq == 0 or counts[t] < k:
            continue
        wf = torch.zeros(model.n_bins, dtype=torch.float64, device="cuda")
        nbr = torch.zeros((nq, k), dtype=torch.int32, device="cuda")
        rad = torch.zeros(nq, dtype=torch.float64, device="cuda")
        runtime.gather(ds, maps[t], t, qs[t], nq, wf, model.gate(), nbr_out=nbr, radius_out=rad)
        torch.cuda.synchronize()
        mx, my, mz = (maps[t].tensors[f][:counts[t]].cpu().numpy() for f in ("x", "y", "z"))
        qpos = np.stack([qs[t].tensors[f][:nq].cpu().numpy() for f in ("x", "y", "z")], 1)
        idx, d2 = oracle.knn(mx, my, mz, qpos, k)
        assert np.array_equal(nbr.cpu().numpy(), idx), t
        assert np.array_equal(rad.cpu().numpy(), np.sqrt(d2[:, -1])), t
        checked += nq
    assert checked >= 900


def test_replay_1e3_slots_1e5_uniforms(gpu):
    model = build_scene(single_shot())
    rs = np.random.default_rng(778)
    replay = rs.random((1000, 100_000))
    _compare(model, (0, 0), abi.PASS_FORWARD, 1000, replay=replay)
