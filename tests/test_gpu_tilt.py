"""Tilted seafloor (an extension of the reference's flat floor) on the GPU vs
the oracle: trajectory parity forward and backward (the tolerance of
test_gpu_trace_parity), the gather of the BOTTOM map with the tilted
multiplier, the whole path, and the batched form with tilted and flat scenes
mixed (bit for bit against the one-scene traces), and the closed-form water
legs of a single tilted layer against the stepped legs of the same layer
split in three.  The oracle's tilted floor is pinned in test_oracle_tilt.py."""
import numpy as np
import pytest

import oracle
from pyalbsim_amd import abi, runtime
from pyalbsim_amd.execution.linear import simulate
from pyalbsim_amd.model import build_scene
from tests.scenes import single_shot, stacked_copies, tilted_shot
from tests.test_gpu_trace_parity import _compare, _download, _match

pytestmark = pytest.mark.gpu

F = ("x", "y", "z", "dx", "dy", "dz", "energy", "time")


@pytest.mark.parametrize("pass_", [abi.PASS_FORWARD, abi.PASS_BACKWARD])
def test_tilted_trace_matches_oracle(gpu, pass_):
    counts = _compare(build_scene(tilted_shot(slope=12.0, azimuth=70.0, b=0.5)), (61, 62), pass_, 400)
    assert counts[abi.BOTTOM_REFLECTION] > 50


@pytest.mark.parametrize("pass_", [abi.PASS_FORWARD, abi.PASS_BACKWARD])
def test_closed_form_legs_match_stepped_legs(gpu, pass_):
    """A single constant layer over a tilted plane keeps closed-form water
    legs; the same layer split in three takes the stepped legs."""
    cfg = tilted_shot(slope=12.0, azimuth=70.0, b=0.6)
    one = runtime.DeviceScene(build_scene(cfg))
    many = runtime.DeviceScene(build_scene(stacked_copies(cfg, 3)))
    assert one.model.tilted and not one.model.layers and many.model.layers
    a, ca = runtime.trace(one, (8, 9), pass_, 2000, tags=True)
    b, cb = runtime.trace(many, (8, 9), pass_, 2000, tags=True)
    assert ca[abi.BOTTOM_REFLECTION] > 200
    bad, allp, n = _match(_download(b, cb), _download(a, ca))
    assert n > 2000
    assert 1.0 - len(bad) / len(allp) >= 0.995


def test_tilted_whole_path_matches_oracle(gpu):
    model = build_scene(tilted_shot(slope=8.0, azimuth=200.0, b=0.4))
    key = (1357, 2468)
    got = simulate(model, 3000, 600, key).waveforms.cpu().numpy()
    ref = oracle.run(model, 3000, 600, key)
    for t in range(4):
        g, r = got[t], ref["waveforms"][t]
        tot = float(np.sum(r))
        if tot == 0:
            assert float(np.sum(g)) == 0
            continue
        assert float(np.sum(g)) == pytest.approx(tot, rel=1e-6), t
        agree = np.abs(g - r) <= 1e-6 * np.maximum(np.abs(r), 1e-300)
        assert np.sum(r[agree]) >= 0.99 * tot, t
    assert float(np.sum(ref["waveforms"][abi.BOTTOM_REFLECTION])) > 0


def test_batched_mixes_tilted_and_flat_scenes(gpu):
    models = [build_scene(tilted_shot(slope=5.0, azimuth=10.0)), build_scene(single_shot()),
              build_scene(tilted_shot(slope=14.0, azimuth=300.0, b=0.8))]
    key = (808, 909)
    counts = [3000, 2500, 2000]
    tab = runtime.DeviceSceneTable(models)
    for pass_ in (abi.PASS_FORWARD, abi.PASS_BACKWARD):
        sinks, tot = runtime.trace_batched(tab, key, pass_, counts, tags=True)
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
