"""Host-side gradient assembly of the adjoint (no GPU): the per-layer formula
of layered columns reduces to the single-layer one, its pieces are the
derivatives of the path log-weight they claim to be, and the Churnside
derivatives per layer equal the single-layer ones for a one-layer column."""
import numpy as np

from pyalbsim_amd import adjoint
from pyalbsim_amd.model import build_scene
from pyalbsim_amd.sweep import shot_config


def _random_sums(rng, n):
    s = rng.uniform(0.5, 3.0, size=(n, adjoint.REPLAY_SUMS))
    s[:, 5] = s[:, 0]  # one layer: every omega0 factor has its free-path density
    return s


def test_layer_formula_equals_single_layer_formula():
    rng = np.random.default_rng(3)
    th = np.array([[0.1, 0.6, 0.9, 5.0, 0.2, 0.3], [0.25, 1.2, 0.85, 12.0, 0.1, 0.4]])
    models = [build_scene(shot_config(t[3], t[0], t[1], t[4], t[5]), hg_g=t[2]) for t in th]
    acc = _random_sums(rng, 2)
    direct = rng.uniform(-1.0, 1.0, size=(2, 3))
    ref = adjoint.gradients_from_sums(acc, direct, [adjoint.scene_derivatives(m, t) for m, t in zip(models, th)])
    for s, m in enumerate(models):
        (lay,) = [(float(m.alpha_l), float(m.absorption), *adjoint._single_layer_dalpha(m))]
        g = adjoint.layer_gradients_from_sums([acc[s]], direct[s], [lay], m.floor_albedo, m.roughness)
        assert np.allclose([g["a"][0], g["b"][0], g["albedo"], g["roughness"]], ref[s, [0, 1, 4, 5]],
                           rtol=1e-7, atol=0)


def test_layer_formula_is_the_log_weight_derivative():
    """For one path prefix with n omega0 factors, d completed free paths, a
    scatter record's own density and water length L, the log-weight is
    (d + [scatter]) log(alpha) + n log((alpha - a) / alpha) - alpha L; the
    formula's alpha- and a-derivatives equal its central differences."""
    n, d, sc, L = 3.0, 3.0, 1.0, 2.7
    alpha, a = 0.45, 0.12

    def logw(al, aa):
        return (d + sc) * np.log(al) + n * np.log((al - aa) / al) - al * L

    sums = np.array([n, sc, L, 0.0, 0.0, d])
    h = 1e-6
    g = adjoint.layer_gradients_from_sums([sums], np.zeros(3), [(alpha, a, 1.0, 0.0)], 1.0, 1.0)
    da_num = (logw(alpha, a + h) - logw(alpha, a - h)) / (2 * h)
    dal_num = (logw(alpha + h, a) - logw(alpha - h, a)) / (2 * h)
    # with d alpha / d a = 1 the a-gradient is the total derivative along alpha = alpha(a)
    assert np.isclose(g["a"][0], dal_num + da_num, rtol=1e-7)
    g_b = adjoint.layer_gradients_from_sums([sums], np.zeros(3), [(alpha, a, 0.0, 1.0)], 1.0, 1.0)
    assert np.isclose(g_b["b"][0], dal_num, rtol=1e-7)


def test_layer_derivatives_of_a_single_layer_column():
    from tests.scenes import stacked_copies

    cfg = shot_config(6.0, 0.15, 0.8, 0.2, 0.3)
    one = build_scene(cfg)
    three = build_scene(stacked_copies(cfg, 3))
    d1 = (float(one.alpha_l), float(one.absorption), *adjoint._single_layer_dalpha(one))
    for d3 in adjoint.layer_derivatives(three):
        assert np.allclose(d3, d1, rtol=1e-9)
