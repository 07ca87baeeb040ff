"""One Adam iteration of the batched inversion over all 4096 synthetic scenes
(depth U[2, 40] m, b U[0.1, 1.5] /m, albedo U[0.05, 0.4]; floor tilted U[0,
15] degrees), at reduced photon counts (2000 forward + 500 backward per scene
instead of 2e5): the batched forward pass plus the path-replay adjoint through
the torch.autograd.Function gives a finite loss and finite gradients for
every scene, a non-zero gradient for most, and the step keeps every estimate
inside its range.  The scenes run in groups of 1024 (forward and backward per
group), each group on its own Philox scene streams.
"""
import numpy as np
import pytest

from pyalbsim_amd import adjoint
from pyalbsim_amd.inversion import (COLS, RANGES, SAMPLE_RATE, UNKNOWN, invert, inversion_scenes, inversion_tilts,
                                    log_waveform_loss, waveform_loss)

pytestmark = pytest.mark.gpu

N = 4096
NF, NB, WL = 2000, 500, 1000
G = 1024  # scenes per batched group


def test_one_iteration_over_4096_scenes(gpu):
    torch = gpu
    th = inversion_scenes(N, seed=3)
    tilts = inversion_tilts(N, seed=3)
    spec = dict(win_len=WL, sample_rate=SAMPLE_RATE, tilts=tilts)
    with torch.no_grad():
        target = adjoint.gated_waveforms(th, NF, NB, (0x7A26, 6), **spec)
    assert target.shape == (N, 4, WL)
    start = th.copy()
    for c in UNKNOWN:
        lo, hi = RANGES[c]
        start[:, COLS[c]] = 0.5 * (lo + hi)
    theta = torch.tensor(start, dtype=torch.float64, device="cuda", requires_grad=True)
    for a in range(0, N, G):  # forward and backward one scene group at a time, as invert does
        b = min(a + G, N)
        sg = dict(spec, tilts=tilts[a:b], scene_base=a)
        w = adjoint.batched_waveforms(theta[a:b], NF, NB, (0x1A7E, 5), depth_rel=1e-2, **sg)
        per = waveform_loss(w, target[a:b]) + log_waveform_loss(w, target[a:b])
        assert bool(torch.isfinite(per).all())
        per.sum().backward()
        del w, per
    g = theta.grad.cpu().numpy()[:, [COLS[c] for c in UNKNOWN]]
    assert np.all(np.isfinite(g))
    assert np.mean(np.any(g != 0.0, axis=1)) > 0.9


def test_invert_one_iteration_keeps_estimates_in_range(gpu):
    th = inversion_scenes(N, seed=3)
    res = invert(th, NF, NB, iterations=1, lr=1e-2, tilts=inversion_tilts(N, seed=3), group_scenes=G)
    assert len(res.depth_rmse) == 2 and np.all(np.isfinite(res.depth_rmse))
    assert np.isfinite(res.loss[0]) and res.scene_loss.shape == (N,) and np.all(np.isfinite(res.scene_loss))
    for c in UNKNOWN:
        v = res.theta[:, COLS[c]]
        lo, hi = RANGES[c]
        assert np.all((v >= lo - 1e-12) & (v <= hi + 1e-12)), c
    assert res.seconds_per_iteration[0] > 0
