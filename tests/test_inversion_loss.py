"""Inversion losses on the host (no GPU): the log-waveform term against a
NumPy restatement, zero at the target, and the hybrid sum."""
import numpy as np
import torch

from pyalbsim_amd.inversion import log_waveform_loss, waveform_loss


def _numpy_log_loss(w, t, floor_rel=1e-4, pool=4):
    s, tt = w.sum(1), t.sum(1)
    L = s.shape[1] // pool * pool
    sp = s[:, :L].reshape(s.shape[0], -1, pool).sum(2)
    tp = tt[:, :L].reshape(tt.shape[0], -1, pool).sum(2)
    eps = floor_rel * tp.max(1, keepdims=True)
    return ((np.log(sp + eps) - np.log(tp + eps)) ** 2).mean(1)


def test_log_waveform_loss():
    rs = np.random.default_rng(0)
    t = rs.exponential(1.0, (3, 4, 1000)) * np.exp(-np.arange(1000) / 150.0)
    w = t * rs.uniform(0.5, 1.5, t.shape)
    got = log_waveform_loss(torch.tensor(w), torch.tensor(t)).numpy()
    assert np.allclose(got, _numpy_log_loss(w, t), rtol=1e-12)
    assert np.all(log_waveform_loss(torch.tensor(t), torch.tensor(t)).numpy() == 0.0)
    # energy moved between bins of one pool of 4, or between photon types, leaves it unchanged
    w3 = t.copy()
    w3[:, 0, 700], w3[:, 0, 701] = w3[:, 0, 700] + 0.5 * w3[:, 0, 701], 0.5 * w3[:, 0, 701]
    w3[:, 1, 10] += w3[:, 2, 10]
    w3[:, 2, 10] = 0.0
    assert np.allclose(log_waveform_loss(torch.tensor(w3), torch.tensor(t)).numpy(), 0.0, atol=1e-24)
    # a late peak 1e-3 of the energy (a faint floor return) shows in both terms of the hybrid loss
    w2 = t.copy()
    w2[:, 0, 700:704] += 1e-3 * t.sum() / 3
    assert np.all(log_waveform_loss(torch.tensor(w2), torch.tensor(t)).numpy() > 0.01)
    assert np.all(waveform_loss(torch.tensor(w2), torch.tensor(t)).numpy() > 0.001)
