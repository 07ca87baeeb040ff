"""Host logic of the contribution kernels' LDS window (no GPU): the gate is
widened to LDS_WINDOW bins, never past the waveform, never narrowed."""
from pyalbsim_amd import runtime


def test_lds_window():
    w = runtime.LDS_WINDOW
    assert runtime.lds_window(40526, 26620, 1024) == (26620, w)
    assert runtime.lds_window(40526, 38000, 1024) == (38000, 40526 - 38000)
    assert runtime.lds_window(40526, 100, 6000) == (100, max(6000, w))
    assert runtime.lds_window(8000, 5300, 1000, runtime.LDS_WINDOW_BATCHED) == (5300, 2700)
    assert runtime.lds_window(100000, 10, 10, runtime.LDS_WINDOW_BATCHED) == (10, runtime.LDS_WINDOW_BATCHED)
