"""alb_gather_static (the gather that allocates nothing) against alb_gather.

Both are called through the C ABI on the same map and queries.  Neighbour
sets and k-th distances must be identical; waveforms may differ only by the
order of FP64 additions (1e-12 of the peak).  Two calls in a row, the second
on a smaller query range, show that the shared static scratch keeps no state
that changes a later result.  The non-direct form (no neighbour arrays) runs
in several rounds of its neighbour-row buffer.
"""
import ctypes as C

import numpy as np
import pytest

from pyalbsim_amd import abi, runtime
from pyalbsim_amd.model import build_scene
from tests.scenes import single_shot

pytestmark = pytest.mark.gpu


def _setup(n_f, n_b, key):
    model = build_scene(single_shot())
    ds = runtime.DeviceScene(model)
    recs, rc = runtime.trace(ds, key, abi.PASS_FORWARD, n_f)
    maps = [runtime.build_map(recs[t], rc[t], runtime.MAP_DIMS[t]) for t in range(4)]
    qs, qc = runtime.trace(ds, key, abi.PASS_BACKWARD, n_b)
    return model, ds, maps, qs, qc


def _with_workspace(torch, L, ds, dmap, t, q, q_begin, q_n, window, direct):
    k = ds.model.k
    wf = torch.zeros(ds.model.n_bins, dtype=torch.float64, device="cuda")
    nbr = torch.full((q_n, k), -7, dtype=torch.int32, device="cuda") if direct else None
    rad = torch.zeros(q_n, dtype=torch.float64, device="cuda") if direct else None
    b = C.c_size_t(0)
    assert L.alb_gather_workspace_size(q_n, k, C.byref(b)) == abi.ALB_OK
    ws = torch.empty(b.value, dtype=torch.uint8, device="cuda")
    rc = L.alb_gather(C.byref(ds.s), C.byref(dmap.struct), t, C.byref(q.struct), q_begin, q_n, runtime._p(wf),
                      ds.model.n_bins, window[0], window[1], runtime._p(nbr), runtime._p(rad), runtime._p(ws),
                      b.value, None, C.c_void_p(runtime._stream_ptr()))
    assert rc == abi.ALB_OK, L.alb_last_error()
    torch.cuda.synchronize()
    return wf, nbr, rad


def _static(torch, L, ds, dmap, t, q, q_begin, q_n, window, direct):
    k = ds.model.k
    wf = torch.zeros(ds.model.n_bins, dtype=torch.float64, device="cuda")
    nbr = torch.full((q_n, k), -7, dtype=torch.int32, device="cuda") if direct else None
    rad = torch.zeros(q_n, dtype=torch.float64, device="cuda") if direct else None
    rc = L.alb_gather_static(C.byref(ds.s), C.byref(dmap.struct), t, C.byref(q.struct), q_begin, q_n,
                             runtime._p(wf), ds.model.n_bins, window[0], window[1], runtime._p(nbr), runtime._p(rad),
                             None, C.c_void_p(runtime._stream_ptr()))
    assert rc == abi.ALB_OK, L.alb_last_error()
    torch.cuda.synchronize()
    return wf, nbr, rad


@pytest.mark.parametrize("direct", [True, False])
def test_static_gather_equals_workspace_gather_on_two_calls(gpu, direct):
    torch = gpu
    L = runtime.load_library()
    model, ds, maps, qs, qc = _setup(60000, 12000, (31, 41))
    window = runtime.lds_window(model.n_bins, *model.gate())
    for t in (abi.SCATTER, abi.BOTTOM_REFLECTION):
        # first call: many queries (several neighbour-row rounds when not direct);
        # second call: a smaller, shifted range
        for q_begin, q_n in ((0, min(qc[t], 40000)), (17, min(qc[t] - 17, 2500))):
            print(f"no-alloc gather: type {t}, direct {direct}, queries [{q_begin}, {q_begin + q_n})", flush=True)
            ref = _with_workspace(torch, L, ds, maps[t], t, qs[t], q_begin, q_n, window, direct)
            got = _static(torch, L, ds, maps[t], t, qs[t], q_begin, q_n, window, direct)
            wr, wg = ref[0].cpu().numpy(), got[0].cpu().numpy()
            assert wr.max() > 0
            assert np.max(np.abs(wg - wr)) <= 1e-12 * np.max(np.abs(wr)), (t, q_begin, q_n)
            if direct:
                assert np.array_equal(ref[1].cpu().numpy(), got[1].cpu().numpy())
                assert np.array_equal(ref[2].cpu().numpy(), got[2].cpu().numpy())


def test_static_gather_rejects_calls_beyond_its_scratch(gpu):
    """Checked on the host before any launch: more queries than the maximum,
    or k beyond the supported 448."""
    torch = gpu
    L = runtime.load_library()
    model, ds, maps, qs, qc = _setup(20000, 100, (5, 6))
    t = abi.SCATTER
    wf = torch.zeros(model.n_bins, dtype=torch.float64, device="cuda")
    n_over = abi.GATHER_STATIC_MAX_QUERIES + 1
    rc = L.alb_gather_static(C.byref(ds.s), C.byref(maps[t].struct), t, C.byref(qs[t].struct), 0, n_over,
                             runtime._p(wf), model.n_bins, 0, 0, None, None, None, C.c_void_p(runtime._stream_ptr()))
    assert rc == abi.ALB_ERR_ARG and b"maximum" in L.alb_last_error()
    s2 = abi.Scene.from_buffer_copy(ds.s)
    s2.k = 449
    rc = L.alb_gather_static(C.byref(s2), C.byref(maps[t].struct), t, C.byref(qs[t].struct), 0, 10,
                             runtime._p(wf), model.n_bins, 0, 0, None, None, None, C.c_void_p(runtime._stream_ptr()))
    assert rc in (abi.ALB_ERR_UNSUPPORTED, abi.ALB_ERR_ARG)
    torch.cuda.synchronize()
    assert float(wf.abs().sum().item()) == 0.0
