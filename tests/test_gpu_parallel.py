"""run_parallel (execution/parallel.py:186-284 of the reference) against the
CPU oracle at the same Philox key.

* The returned PhotonMapData hold the oracle's forward records: every oracle
  record has a GPU record at the same sub-step whose position (per axis,
  1e-5 of max(|coordinate|, 1 m)) and energy (1e-5 relative) agree, for at
  least 99.5 % of the records (the trace parity tolerance; the maps carry no
  tags, so records are matched by nearest position), and the counts agree to
  0.5 %.
* The float32 waveforms agree with the oracle's FP64 whole-path waveforms:
  total energy to 1e-5 and at least 99 % of the energy in bins that agree to
  1e-5.
* same_seed_batches=True reproduces the reference's seeding (every batch the
  same photons): the maps hold batches_forward copies of one batch's records
  and the waveform equals batches_backward times the oracle's gather of one
  backward batch over those repeated records.
"""
import numpy as np
import pytest
from scipy.spatial import cKDTree

import oracle
from pyalbsim_amd import abi
from pyalbsim_amd.model import build_scene
from pyalbsim_amd.photon_mapping.photon_type import PhotonType
from tests.scenes import single_shot

pytestmark = pytest.mark.gpu


def _match_records(gpu_map, orc, tol=1e-5):
    """Fraction of oracle records with a matching GPU record."""
    if len(orc["x"]) == 0:
        return 1.0
    gp = gpu_map.positions.astype(np.float64)
    op = np.stack([orc["x"], orc["y"], orc["z"]], 1).astype(np.float64)
    d, j = cKDTree(gp).query(op, k=1)
    scale = np.maximum(np.abs(op), 1.0)
    ok = np.all(np.abs(gp[j] - op) <= tol * scale, axis=1)
    ok &= np.abs(gpu_map.times[j].astype(np.float64) - orc["time"]) <= 1.0
    ok &= np.abs(gpu_map.energies[j] - orc["energy"]) <= tol * np.abs(orc["energy"])
    return float(np.mean(ok))


def _waveforms_agree(got, ref, tol=1e-5):
    tot_g, tot_r = float(np.sum(got)), float(np.sum(ref))
    assert abs(tot_g - tot_r) <= tol * abs(tot_r), (tot_g, tot_r)
    good = np.abs(got - ref) <= tol * np.abs(ref) + 1e-30
    assert float(np.sum(np.abs(ref)[good])) >= 0.99 * float(np.sum(np.abs(ref)))


def _rc(bf, nf, bb, nb):
    from alb_sim.config.run import RunConfig

    return RunConfig(photons_per_batch_forward=nf, batches_forward=bf, photons_per_batch_backward=nb,
                     batches_backward=bb, processes=4)


def test_run_parallel_matches_the_oracle(gpu):
    from alb_sim.execution.parallel import run_parallel

    cfg = single_shot()
    model = build_scene(cfg)
    rc = _rc(2, 1500, 2, 400)
    seed = 23
    wf, maps, _ = run_parallel(cfg, rc, seed=seed, verbose=False)
    key = abi.key_from_generator(np.random.default_rng(seed))
    ref = oracle.run(model, 3000, 800, key)
    for pt in PhotonType:
        t = pt.abi_index
        n_g, n_o = len(maps[pt]), len(ref["maps"][t]["x"])
        assert abs(n_g - n_o) <= max(2, 0.005 * n_o), (pt, n_g, n_o)
        assert _match_records(maps[pt], ref["maps"][t]) >= 0.995, pt
        _waveforms_agree(wf[pt].astype(np.float64), ref["waveforms"][t])


def test_same_seed_batches_reproduce_the_reference_seeding(gpu):
    from alb_sim.execution.parallel import run_parallel

    cfg = single_shot()
    model = build_scene(cfg)
    bf, nf, bb, nb = 4, 1000, 3, 300
    seed = 5
    wf, maps, _ = run_parallel(cfg, _rc(bf, nf, bb, nb), seed=seed, verbose=False, same_seed_batches=True)
    key = abi.key_from_generator(np.random.default_rng(seed))
    one, _, _ = oracle.trace(model, key, abi.PASS_FORWARD, 0, nf)
    qs, _, _ = oracle.trace(model, key, abi.PASS_BACKWARD, 0, nb)
    for pt in PhotonType:
        t = pt.abi_index
        m = maps[pt]
        assert len(m) % bf == 0
        # every record appears bf times
        u, c = np.unique(m.positions, axis=0, return_counts=True)
        assert np.all(c % bf == 0), pt
        rep = {f: np.tile(v, bf) for f, v in one[t].items()}
        assert abs(len(m) - len(rep["x"])) <= bf * 2
        if len(qs[t]["x"]) == 0:
            assert float(np.sum(wf[pt])) == 0.0
            continue
        ref = bb * oracle.gather(model, t, rep, qs[t])
        _waveforms_agree(wf[pt].astype(np.float64), ref)
