"""The 1024-scene Latin-hypercube sweep (seed 1) through run_sweep, at reduced
photon counts (10^U[3, 4] per scene instead of 10^U[5, 7], backward = forward
/ 4), the batched form of the notebook's per-scene run_linear loop
(notebooks/run_hypercube_sampling.ipynb cells 4, 10):

* every scene's gated waveform is finite and non-negative, its surface
  return lies in the first bins of the gate, and every scene has signal;
* sampled scenes equal the one-scene path (simulate with the scene's stream)
  up to FP64 summation order, and the waveforms do not depend on how the
  sweep is split into memory groups;
* two scenes against the CPU oracle over the same streams: the whole path's
  waveforms agree (total energy to 1e-6, 99 % of the energy in bins that
  agree to 1e-6).
"""
import numpy as np
import pytest

import oracle
from pyalbsim_amd import abi
from pyalbsim_amd.execution.linear import simulate
from pyalbsim_amd.sweep import hypercube, run_sweep

pytestmark = pytest.mark.gpu

KEY = (0xBA7C, 1)
WL = 1000


@pytest.fixture(scope="module")
def sweep(gpu):
    sc = hypercube(1024, 1, (3.0, 4.0))
    res = run_sweep(sc, KEY, win_len=WL, strict=False)
    return sc, res


def test_sweep_of_1024_scenes_is_sound(sweep):
    sc, res = sweep
    assert len(res.waveforms) == 1024
    for s, (x, w) in enumerate(zip(sc, res.waveforms)):
        m = x.model
        assert w.shape == (4, m.gate(WL)[1])
        assert np.all(np.isfinite(w)) and np.all(w >= 0), s
        assert w.sum() > 0, s
        sb = m.surface_return_bin() - res.windows[s][0]
        assert 0 <= sb < 200, (s, sb)
        head = w[abi.SURFACE_REFLECTION, :int(sb) + 20].sum()
        assert head > 0, s  # the surface return shows at its bin


@pytest.mark.parametrize("s", [0, 17, 511, 1023])
def test_sweep_scenes_equal_one_scene_runs(sweep, s):
    sc, res = sweep
    x = sc[s]
    one = simulate(x.model, x.n_forward, x.n_backward, KEY, scene_id=s, strict=False)
    g0, gl = res.windows[s]
    ref = one.waveforms[:, g0:g0 + gl].cpu().numpy()
    got = res.waveforms[s]
    assert np.max(np.abs(got - ref)) <= 1e-9 * np.max(np.abs(ref)), s


def test_sweep_grouping_does_not_change_waveforms(sweep):
    sc, res = sweep
    part = sc[:96]
    small = run_sweep(part, KEY, win_len=WL, strict=False, budget_bytes=2.0e9)
    assert len(small.groups) > 1
    for s in range(96):
        a, b = small.waveforms[s], res.waveforms[s]
        assert np.max(np.abs(a - b)) <= 1e-9 * max(np.max(np.abs(b)), 1e-300), s


def test_two_sweep_scenes_match_the_oracle(sweep):
    sc, res = sweep
    order = np.argsort([x.n_forward for x in sc])
    for s in (int(order[0]), int(order[1])):  # the two smallest scenes
        x = sc[s]
        m = x.model
        maps, _, _ = oracle.trace(m, KEY, abi.PASS_FORWARD, 0, x.n_forward, stream=abi.PASS_FORWARD | (s << 1))
        qs, _, _ = oracle.trace(m, KEY, abi.PASS_BACKWARD, 0, x.n_backward, stream=abi.PASS_BACKWARD | (s << 1))
        g0, gl = res.windows[s]
        got = res.waveforms[s]
        for t in range(4):
            if len(qs[t]["x"]) == 0 or len(maps[t]["x"]) < m.k:
                assert got[t].sum() == 0.0
                continue
            ref = oracle.gather(m, t, maps[t], qs[t])[g0:g0 + gl]
            tot_g, tot_r = float(got[t].sum()), float(ref.sum())
            assert abs(tot_g - tot_r) <= 1e-6 * abs(tot_r), (s, t, tot_g, tot_r)
            good = np.abs(got[t] - ref) <= 1e-6 * np.abs(ref) + 1e-300
            assert float(ref[good].sum()) >= 0.99 * tot_r, (s, t)
