"""Depth start of the inversion from the floor return's leading edge, on the
host (no GPU), with synthetic waveforms whose floor arrival is known:

* an exponentially decaying volume return plus a floor pulse starting at the
  bin 2 D / (v_water x bin width) after the surface gives D back to within two
  pooled bins, for several depths, bin scales and surface positions (a wrong
  sign or origin in the bin -> depth conversion fails these);
* a pure volume decay, with or without a floor pulse below floor_rel x peak,
  gives NaN (no start claimed), as does an empty waveform;
* the estimate is the same for the waveform split over photon types and summed
  by the caller, and for any positive scale of the waveform.
"""
import numpy as np
import pytest

from pyalbsim_amd.inversion import floor_timing_depth


def _waveform(L, surface_bin, bins_per_m, depth, floor_amp, decay_per_bin=0.02, noise=0.0, seed=0):
    t = np.arange(L, dtype=np.float64)
    w = np.where(t >= surface_bin, np.exp(-decay_per_bin * (t - surface_bin)), 0.0)
    w[int(surface_bin):int(surface_bin) + 3] += 50.0  # surface return
    if depth is not None:
        t_f = surface_bin + depth * bins_per_m
        pulse = np.where(t >= t_f, np.exp(-(t - t_f) / 6.0), 0.0)
        w += floor_amp * np.exp(-decay_per_bin * depth * bins_per_m) * pulse
    if noise:
        w *= np.random.default_rng(seed).lognormal(0.0, noise, L)
    return w


@pytest.mark.parametrize("bpm,sb", [(17.879, 64.9), (8.94, 30.2), (35.76, 120.0)])
@pytest.mark.parametrize("depth", [3.0, 9.5, 21.0, 37.0])
def test_floor_arrival_is_recovered(bpm, sb, depth):
    L = int(sb + 45 * bpm)
    w = _waveform(L, sb, bpm, depth, floor_amp=30.0, decay_per_bin=0.3 / bpm, noise=0.05, seed=int(depth))
    d = floor_timing_depth(w[None], bpm, sb)
    assert np.isfinite(d[0])
    assert abs(d[0] - depth) <= 8.0 / bpm + 1e-9, (d[0], depth)  # two pooled bins


def test_no_floor_return_gives_nan():
    bpm, sb, L = 17.879, 64.9, 1000
    plain = _waveform(L, sb, bpm, None, 0.0, noise=0.05, seed=1)
    buried = _waveform(L, sb, bpm, 30.0, 1e-9, decay_per_bin=0.05, noise=0.0)  # below floor_rel x peak
    empty = np.zeros(L)
    d = floor_timing_depth(np.stack([plain, buried, empty]), bpm, sb)
    assert np.all(np.isnan(d)), d


def test_per_scene_parameters_scale_and_type_split():
    rs = np.random.default_rng(3)
    bpm = np.array([17.879, 8.94, 17.879])
    sb = np.array([64.9, 30.2, 80.0])
    depth = np.array([12.0, 30.0, 4.0])
    L = 1000
    w = np.stack([_waveform(L, sb[i], bpm[i], depth[i], 30.0, decay_per_bin=0.3 / bpm[i]) for i in range(3)])
    d = floor_timing_depth(w, bpm, sb)
    assert np.all(np.abs(d - depth) <= 8.0 / bpm + 1e-9), d
    split = rs.dirichlet(np.ones(4), size=(3, L)).transpose(0, 2, 1) * w[:, None, :]  # [S, 4, L]
    assert np.allclose(floor_timing_depth(split.sum(axis=1), bpm, sb), d)
    assert np.array_equal(floor_timing_depth(w * 7.5e-3, bpm, sb), d)


def test_unknown_start_is_refused_before_any_device_work():
    from pyalbsim_amd.inversion import invert, inversion_scenes

    with pytest.raises(ValueError, match="start"):
        invert(inversion_scenes(2), 100, 25, iterations=1, start="grid")


def _counted(L, surface_bin, bins_per_m, n0, decay_per_bin, sigma_w, rs, depth=None, floor_amp=0.0):
    """Monte Carlo-like waveform: each raw bin holds Poisson(lambda) photons of
    lognormal(0, sigma_w) weight; lambda is the expected volume decay (plus an
    optional floor pulse), n0 photons per bin at the surface."""
    t = np.arange(L, dtype=np.float64)
    lam = np.where(t >= surface_bin, n0 * np.exp(-decay_per_bin * (t - surface_bin)), 0.0)
    lam[int(surface_bin):int(surface_bin) + 3] += 50.0 * n0
    if depth is not None:
        t_f = surface_bin + depth * bins_per_m
        lam += floor_amp * n0 * np.exp(-decay_per_bin * depth * bins_per_m) * np.where(
            t >= t_f, np.exp(-(t - t_f) / 6.0), 0.0)
    c = rs.poisson(lam)
    w = np.zeros(L)
    # exact sums of few weights; the normal limit of the sum for many
    m1 = np.exp(sigma_w ** 2 / 2)
    sd1 = np.sqrt((np.exp(sigma_w ** 2) - 1) * np.exp(sigma_w ** 2))
    few = (c > 0) & (c <= 64)
    for b in np.nonzero(few)[0]:
        w[b] = rs.lognormal(0.0, sigma_w, c[b]).sum()
    many = c > 64
    w[many] = c[many] * m1 + np.sqrt(c[many]) * sd1 * rs.standard_normal(np.count_nonzero(many))
    return w


@pytest.mark.parametrize("n0,decay,sigma_w", [(1e4, 0.3, 0.5), (1e5, 0.3, 0.3), (1e4, 0.15, 0.5), (1e3, 0.3, 0.5)])
def test_sparse_volume_tail_without_floor_gives_nan(n0, decay, sigma_w):
    """A volume return that runs out of photons (Poisson counts, lognormal
    weights, no floor) must not be read as a floor return."""
    bpm, sb, L = 17.879, 64.9, 1000
    rs = np.random.default_rng(int(n0) + int(100 * decay))
    w = np.stack([_counted(L, sb, bpm, n0, decay / bpm, sigma_w, rs) for _ in range(200)])
    d = floor_timing_depth(w, bpm, sb)
    assert np.count_nonzero(np.isfinite(d)) <= 2, np.sort(d[np.isfinite(d)])


@pytest.mark.parametrize("depth", [6.0, 15.0, 25.0])
def test_counted_floor_return_is_recovered(depth):
    """The same counting noise with a floor pulse that carries enough photons."""
    bpm, sb, L = 17.879, 64.9, 1000
    rs = np.random.default_rng(int(depth))
    w = np.stack([_counted(L, sb, bpm, 1e4, 0.3 / bpm, 0.5, rs, depth, floor_amp=30.0) for _ in range(20)])
    d = floor_timing_depth(w, bpm, sb)
    assert np.all(np.abs(d - depth) <= 8.0 / bpm + 1e-9), d
