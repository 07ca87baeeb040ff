"""Pins of the Henyey-Greenstein phase table and backscatter fraction and of
the Cox-Munk roughness (the sweep / adjoint / inversion scenes' extensions of
the reference's Fournier-Forand table, fournier_forand_phase_function.py:9-48,
and its roughness field, config/sea_surface.py:12-14).

Each check comes from a property the mathematics fixes, not from the formula
in pyalbsim_amd.model: the solid-angle normalisation, the mean cosine g (the
defining property of the HG asymmetry), the forward/backward ratio
((1+g)/(1-g))^3, the closed-form angular CDF, the isotropic limit, quadrature
of the backward hemisphere, and the Cox-Munk (1954) clean-surface regression
of the mean-square slope on wind speed.
"""
import math

import numpy as np
import pytest
from scipy import integrate

from pyalbsim_amd import abi
from pyalbsim_amd.model import (build_scene, cox_munk_roughness, henyey_greenstein_backscatter,
                                henyey_greenstein_table)
from pyalbsim_amd.sweep import RANGES, hypercube, shot_config

GS = (0.80, 0.85, 0.90, 0.93, 0.95)


def _simpson(y, x):
    return integrate.simpson(y, x=x)


@pytest.mark.parametrize("g", GS)
def test_hg_table_normalised_over_the_sphere(g):
    theta, pf, _ = henyey_greenstein_table(g)
    # 2 pi int p sin dtheta = 1 (the grid starts at 1e-6 rad: the missing cap is
    # p(0) * pi * 1e-12 < 1e-9)
    assert 2 * np.pi * _simpson(pf * np.sin(theta), theta) == pytest.approx(1.0, abs=1e-6)


@pytest.mark.parametrize("g", GS)
def test_hg_mean_cosine_is_g(g):
    theta, pf, cdf = henyey_greenstein_table(g)
    mu = 2 * np.pi * _simpson(pf * np.sin(theta) * np.cos(theta), theta)
    assert mu == pytest.approx(g, abs=1e-6)
    # the sampling table (Riemann-sum CDF) reproduces it up to its bin error:
    # each bin's mass sits at its node, so the mean moves by at most one bin
    # width in theta times |d cos / d theta| <= 1
    dmass = np.diff(np.concatenate([[0.0], cdf]))
    mu_tab = float(np.sum(dmass * np.cos(theta)))
    assert abs(mu_tab - g) <= np.pi / 18000


@pytest.mark.parametrize("g", GS)
def test_hg_forward_backward_ratio(g):
    theta, pf, _ = henyey_greenstein_table(g)
    # p(0) / p(pi) = ((1 + g) / (1 - g))^3; theta[0] = 1e-6 changes p(0) by < 1e-9 relative
    assert pf[0] / pf[-1] == pytest.approx(((1 + g) / (1 - g)) ** 3, rel=1e-8)
    # absolute level at pi: (1 - g^2) / (4 pi (1 + g)^3) = (1 - g) / (4 pi (1 + g)^2)
    assert pf[-1] == pytest.approx((1 - g) / (4 * np.pi * (1 + g) ** 2), rel=1e-12)


@pytest.mark.parametrize("g", GS)
def test_hg_table_cdf_matches_the_closed_form_angular_cdf(g):
    """P(angle <= theta) = (1 - g^2) / (2 g) (1 / (1 - g) - 1 / sqrt(1 + g^2 - 2 g cos theta)):
    the table's Riemann-sum CDF may differ from it by one bin's mass."""
    theta, pf, cdf = henyey_greenstein_table(g)
    F = (1 - g * g) / (2 * g) * (1 / (1 - g) - 1 / np.sqrt(1 + g * g - 2 * g * np.cos(theta)))
    bin_mass = 2 * np.pi * pf * np.sin(theta) * (theta[1] - theta[0])
    assert np.max(np.abs(cdf - F)) <= 1.05 * np.max(bin_mass)
    assert cdf[-1] == 1.0 and np.all(np.diff(cdf) >= 0)
    # inverse-CDF sampling of the table: the sampled mean cosine is g
    u = np.random.default_rng(0).random(400000)
    th = np.interp(u, cdf, theta)
    assert np.mean(np.cos(th)) == pytest.approx(g, abs=3 * np.std(np.cos(th)) / np.sqrt(len(u)) + np.pi / 18000)


def test_hg_isotropic_limit():
    theta, pf, cdf = henyey_greenstein_table(0.0)
    assert np.allclose(pf, 1 / (4 * np.pi), rtol=1e-14)
    assert henyey_greenstein_backscatter(0.0) == 0.5
    assert henyey_greenstein_backscatter(1e-6) == pytest.approx(0.5, abs=1e-6)


@pytest.mark.parametrize("g", GS + (0.3, 0.6))
def test_hg_backscatter_fraction_by_quadrature(g):
    """B = 2 pi int_{pi/2}^{pi} p sin: quadrature of the model's own table
    (Simpson over its nodes) and of the textbook density."""
    theta, pf, _ = henyey_greenstein_table(g)
    back = theta >= np.pi / 2
    # start the Simpson sum at the first node past pi/2 and add the sliver before it
    t0 = theta[back][0]
    sliver = integrate.quad(lambda t: np.interp(t, theta, pf) * np.sin(t), np.pi / 2, t0)[0]
    b_tab = 2 * np.pi * (_simpson(pf[back] * np.sin(theta[back]), theta[back]) + sliver)
    assert henyey_greenstein_backscatter(g) == pytest.approx(b_tab, rel=1e-7)

    def p(t):
        return (1 - g * g) / (4 * math.pi * (1 + g * g - 2 * g * math.cos(t)) ** 1.5)

    b_quad = 2 * math.pi * integrate.quad(lambda t: p(t) * math.sin(t), math.pi / 2, math.pi)[0]
    assert henyey_greenstein_backscatter(g) == pytest.approx(b_quad, rel=1e-10)


def test_hg_scene_uses_the_hg_table_and_backscatter():
    g = 0.9
    m = build_scene(shot_config(5.0, 0.1, 0.2, 0.15, 0.05), hg_g=g)
    theta, _, cdf = henyey_greenstein_table(g)
    assert np.array_equal(m.phase_cdf, cdf) and np.array_equal(m.phase_theta, theta)
    assert m.backscatter == henyey_greenstein_backscatter(g)
    assert m.phase_g == g


@pytest.mark.parametrize("wind", [0.0, 5.0, 10.0])
def test_cox_munk_roughness(wind):
    """Cox and Munk (1954), clean surface: mean-square slope
    sigma^2 = 0.003 + 5.12e-3 W (W in m/s at 12.5 m).  The roughness r is
    taken so that the GGX alpha = r^2 is the RMS slope sigma."""
    r = cox_munk_roughness(wind)
    sigma2 = 0.003 + 5.12e-3 * wind
    assert r ** 4 == pytest.approx(sigma2, rel=1e-13)
    m = build_scene(shot_config(5.0, 0.1, 0.2, 0.15, r))
    assert abi.make_scene(m, 0, 0, 0).alpha_ggx == pytest.approx(math.sqrt(sigma2), rel=1e-13)


def test_cox_munk_published_points_and_sweep_range():
    # RMS slopes from the regression: calm 0.0548, 10 m/s 0.2328
    assert cox_munk_roughness(0.0) ** 2 == pytest.approx(0.05477, abs=1e-5)
    assert cox_munk_roughness(10.0) ** 2 == pytest.approx(0.23281, abs=1e-5)
    # monotone in wind; the sweep's roughness column follows its wind column
    ws = np.linspace(*RANGES["wind"], 11)
    rs = [cox_munk_roughness(w) for w in ws]
    assert np.all(np.diff(rs) > 0)
    for s in hypercube(16, 1, (3.0, 3.0)):
        assert s.params["roughness"] == cox_munk_roughness(s.params["wind"])
        assert abi.make_scene(s.model, 0, 0, 0).alpha_ggx == pytest.approx(s.params["roughness"] ** 2, rel=1e-15)
