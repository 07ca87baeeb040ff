"""Graph-captured batched trace (runtime.BatchedTraceGraph) and graph reuse in
simulate_batched.

Every replay of a captured pass, with eager GPU work between replays, must
equal the eager batched trace of the same scenes bit for bit (records matched
by tag, all fields), forward and backward; after an in-place update of the
scene table the replay must equal the eager trace of the new scenes; and
simulate_batched with a graph cache over several calls (scenes A, B, A) must
give the waveforms of the uncached calls to 1e-9 of the peak (FP64 summation
order).
"""
import numpy as np
import pytest

from pyalbsim_amd import abi, runtime
from pyalbsim_amd.execution.batched import simulate_batched
from pyalbsim_amd.model import build_scene
from tests.scenes import single_shot

pytestmark = pytest.mark.gpu

F = ("x", "y", "z", "dx", "dy", "dz", "energy", "time", "scene")


def _models(shift=0.0):
    return [build_scene(single_shot(depth=5.0 + shift)),
            build_scene(single_shot(depth=12.0 + shift, a=0.05, b=0.9, albedo=0.3, roughness=0.3), hg_g=0.9),
            build_scene(single_shot(depth=2.0 + shift, a=0.3, b=1.5, albedo=0.05, roughness=0.45), hg_g=0.82)]


def _sorted(sinks, counts):
    out = []
    for t in range(4):
        v = sinks[t].view(counts[t])
        tag = v["tag"].cpu().numpy()
        sc = v["scene"].cpu().numpy()
        o = np.lexsort((tag, sc))
        out.append({f: v[f].cpu().numpy()[o] for f in F} | {"tag": tag[o]})
    return out


def _equal(a, b):
    for t in range(4):
        for f in a[t]:
            assert np.array_equal(a[t][f], b[t][f]), (t, f)


@pytest.mark.parametrize("pass_", [abi.PASS_FORWARD, abi.PASS_BACKWARD])
def test_replays_equal_eager_trace(gpu, pass_):
    torch = gpu
    key = (77, 78)
    counts = [3000, 5000, 2000]
    tab = runtime.DeviceSceneTable(_models())
    ref_s, ref_c = runtime.trace_batched(tab, key, pass_, counts, tags=True)
    ref = _sorted(ref_s, ref_c)
    g = runtime.BatchedTraceGraph(tab, key, pass_, counts, [2 * c + 4096 for c in ref_c], tags=True)
    other = runtime.DeviceScene(build_scene(single_shot(depth=9.0)))
    for i in range(3):
        runtime.trace(other, (i, i), abi.PASS_FORWARD, 4000)  # eager work between replays
        torch.cuda.synchronize()
        s, c = g.replay()
        assert c == ref_c
        _equal(_sorted(s, c), ref)
    # new scenes in the same table: the replay traces them
    new = _models(shift=1.5)
    tab.update(new)
    s, c = g.replay()
    fresh = runtime.DeviceSceneTable(new)
    e_s, e_c = runtime.trace_batched(fresh, key, pass_, counts, tags=True)
    assert c == e_c
    _equal(_sorted(s, c), _sorted(e_s, e_c))


def test_simulate_batched_with_graph_cache(gpu):
    key = (91, 92)
    nf, nb = [20000, 8000, 12000], [3000, 2000, 2500]
    cache = {}
    for models in (_models(), _models(shift=2.0), _models()):
        a = simulate_batched(models, nf, nb, key, win_len=1000, graphs=cache)
        b = simulate_batched(models, nf, nb, key, win_len=1000)
        for s in range(3):
            wa, wb = a.waveform(s).cpu().numpy(), b.waveform(s).cpu().numpy()
            for t in range(4):
                peak = float(np.max(np.abs(wb[t])))
                assert np.max(np.abs(wa[t] - wb[t])) <= 1e-9 * peak + 1e-300, (s, t)
        assert list(map(list, a.scene_map_counts)) == list(map(list, b.scene_map_counts))
    assert sum(1 for k in cache if k != "table") == 2  # one graph per pass, replayed
