"""Host logic of the learned multiplier's training (learned.py) on synthetic
step logs: the samples' inputs, targets and weights, and the P(k) error."""
import math

import numpy as np

import learned as LR


class HP:
    Nmesh, Nsample, Box, model, modified_gravity_active, fofr0 = 32, 32, 64.0, 1, 1, 1e-5


def logs(ratio_final):
    nb = 8
    k = np.linspace(0.0, 1.4, nb)
    modes = np.arange(nb, dtype=float) * 10
    a_s = [0.05, 0.5, 1.0]
    fine, coarse = [], []
    for i, a in enumerate(a_s):
        pf = 100.0 * a * np.ones(nb)
        pc = pf / (ratio_final if i == len(a_s) - 1 else 1.0)
        fine.append((a, 0.0, pf, k, modes))
        coarse.append((a, 0.3 + a, pc, k, modes))
    return fine, coarse, k, modes


def test_samples_use_the_output_time_ratio_at_every_step():
    r = np.array([1.0, 1.1, 1.2, 1.3, 1.5, 2.0, 3.0, 9.0])
    fine, coarse, k, modes = logs(r)
    kmax = math.pi * 32 / 64.0
    x, t, w = LR.samples(fine, coarse, HP, None, kmax, eta=0.6, clip=0.3)
    ok = (modes > 0) & (k > 0) & (k < kmax)  # bin 0 (k = 0, no modes) is dropped
    nk = ok.sum()
    assert x.shape == (3 * nk, 4) and t.shape == w.shape == (3 * nk,)
    kny = math.pi * 32 / 64.0
    for s, a in enumerate([0.05, 0.5, 1.0]):
        blk = slice(s * nk, (s + 1) * nk)
        np.testing.assert_allclose(x[blk, 0], k[ok] / kny, rtol=1e-6)
        assert np.all(x[blk, 1] == np.float32(a))
        assert np.all(x[blk, 2] == np.float32((math.log10(1e-5) + 6) / 2))
        np.testing.assert_allclose(x[blk, 3], 0.3 + a, rtol=1e-6)
        np.testing.assert_allclose(t[blk], np.clip(0.6 * np.log(r[ok]), -0.3, 0.3), rtol=1e-6)
        np.testing.assert_array_equal(w[blk], modes[ok])


def test_theta_and_pofk_error():
    class LC(HP):
        modified_gravity_active = 0
    assert LR.theta_of(LC) == -1.0
    assert LR.theta_of(HP) == (math.log10(1e-5) + 6) / 2
    r = np.array([1.0, 1.1, 1.2, 1.3, 1.5, 2.0, 3.0, 9.0])
    fine, coarse, k, _ = logs(r)
    err, ks = LR.pofk_error(fine, coarse, kmax=1.0)
    sel = (k > 0) & (k < 1.0)
    assert np.allclose(ks, k[sel])
    assert abs(err - np.max(np.abs(1 / r[sel] - 1))) < 1e-12


def test_samples_weight_power_and_k0_anchors():
    """wpow sets the weights to modes**wpow; anchor adds, after each step's bins, one sample at k = 0 with
    target ln M = 0 (not a correction of the current network) and anchor times the step's mean weight."""
    r = np.array([1.0, 1.1, 1.2, 1.3, 1.5, 2.0, 3.0, 9.0])
    fine, coarse, k, modes = logs(r)
    kmax = math.pi * 32 / 64.0
    ok = (modes > 0) & (k > 0) & (k < kmax)
    nk = ok.sum()

    class Net:  # a network whose output is 0.7 everywhere: targets are 0.7 + correction, anchors stay 0
        def forward(self, x):
            return np.full(len(x), 0.7, np.float32)

    x, t, w = LR.samples(fine, coarse, HP, Net(), kmax, eta=0.6, clip=0.3, wpow=0.25, anchor=4.0)
    assert x.shape == (3 * (nk + 1), 4)
    for s, a in enumerate([0.05, 0.5, 1.0]):
        blk = slice(s * (nk + 1), s * (nk + 1) + nk)
        anc = s * (nk + 1) + nk
        np.testing.assert_allclose(w[blk], modes[ok] ** 0.25, rtol=1e-6)
        np.testing.assert_allclose(t[blk], 0.7 + np.clip(0.6 * np.log(r[ok]), -0.3, 0.3), rtol=1e-6)
        assert x[anc, 0] == 0.0 and x[anc, 1] == np.float32(a) and t[anc] == 0.0
        np.testing.assert_allclose(x[anc, 3], 0.3 + a, rtol=1e-6)
        np.testing.assert_allclose(w[anc], 4.0 * np.mean(modes[ok] ** 0.25), rtol=1e-6)
