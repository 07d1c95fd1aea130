"""The learned multiplier's network on the GPU (mgpu_mlp_*): forward pass,
gradient and training against the numpy restatement in oracle/mlp.py, the
density feature against the oracle's deposit, and the force pass with the
multiplier table against the same pass on delta(k) M(|n|^2)."""
import mlp as Q
import numpy as np
import pytest

import mgpicola as M
import oracle as O
from helpers import lattice_particles, rel_err

pytestmark = pytest.mark.gpu
OM = 0.3


def random_params(seed, scale=0.4):
    rng = np.random.default_rng(seed)
    p = rng.normal(size=Q.NPARAMS) * scale
    return p.astype(np.float32).astype(np.float64)


def inputs(n, seed):
    rng = np.random.default_rng(seed)
    x = np.stack([rng.uniform(0, 1.7, n), rng.uniform(0.05, 1, n), rng.choice([-1.0, 0.5, -0.5], n),
                  rng.uniform(0.1, 1.5, n)], axis=1)
    return x.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("n", [1, 16, 1003, 40000])
def test_forward_matches_oracle(n):
    """BF16 weights and activations with FP32 accumulation (MFMA) against the
    oracle rounding at the same points in FP64.  An activation whose FP32 and
    FP64 values straddle a BF16 rounding boundary rounds one BF16 step apart
    (about 5e-5 of the 128 rounded activations of a row), so 98 % of the rows
    agree to 1e-4 of the output range and the rest to 2e-2."""
    m = M.MLP(seed=2)
    p = random_params(11)
    m.params = p
    np.testing.assert_array_equal(m.params, p.astype(np.float32))
    x = inputs(n, 5)
    y = m.forward(x).astype(np.float64)
    ref = Q.forward(p, x, bf16_on=True)
    scale = max(np.abs(ref).max(), 1.0)
    d = np.abs(y - ref) / scale
    assert d.max() < 2e-2, d.max()
    assert np.quantile(d, 0.98) < 1e-4, np.quantile(d, 0.98)
    # and the roundings matter: the exact FP64 function is further away
    exact = Q.forward(p, x, bf16_on=False)
    assert np.abs(exact - ref).max() / scale > 10 * np.quantile(d, 0.98) or n < 16
    m.close()


def test_initial_network_is_the_identity_multiplier():
    """The output layer starts at zero: ln M = 0 for every input."""
    m = M.MLP(seed=9)
    p = m.params
    o = Q.unpack(p)
    assert np.all(o["W4"] == 0) and o["b4"][0] == 0
    assert np.all(m.forward(inputs(100, 1)) == 0)
    m.close()


@pytest.mark.parametrize("n", [37, 5000])
def test_gradient_matches_oracle_and_is_deterministic(n):
    m = M.MLP(seed=3)
    p = random_params(21, 0.3)
    m.params = p
    x = inputs(n, 6)
    rng = np.random.default_rng(8)
    t = rng.normal(size=n).astype(np.float32)
    w = rng.uniform(0.2, 2.0, n).astype(np.float32)
    loss, g = m.gradient(x, t, w)
    lref, gref = Q.loss_and_grad(p, x, t.astype(np.float64), w.astype(np.float64), bf16_on=True)
    assert abs(loss - lref) <= 1e-4 * lref
    gu, ru = Q.unpack(g), Q.unpack(gref)
    for name, _ in Q.SIZES:
        s = np.abs(ru[name]).max()
        assert np.abs(gu[name] - ru[name]).max() <= 5e-3 * s, name
    loss2, g2 = m.gradient(x, t, w)
    assert loss2 == loss and np.array_equal(g, g2)
    m.close()


def test_training_fits_a_smooth_function():
    """Full-batch Adam on the GPU fits ln M = 0.3 sin(3 k) a + 0.1 theta s to a
    small fraction of its variance."""
    m = M.MLP(seed=4)
    x = inputs(4000, 12).astype(np.float32)
    t = (0.3 * np.sin(3 * x[:, 0]) * x[:, 1] + 0.1 * x[:, 2] * x[:, 3]).astype(np.float32)
    w = np.ones(len(x), np.float32)
    l0 = m.gradient(x, t, w)[0]
    m.train(x, t, w, 1500, 3e-3)
    l1 = m.gradient(x, t, w)[0]
    assert l1 < 0.02 * l0, (l0, l1)
    # the trained parameters reproduce the oracle's prediction
    y = m.forward(x)
    ref = Q.forward(m.params.astype(np.float64), x.astype(np.float64))
    d = np.abs(y - ref) / max(np.abs(ref).max(), 1.0)
    assert np.quantile(d, 0.98) < 1e-4 and d.max() < 2e-2
    m.close()


def test_density_feature_and_force_multiplier():
    """s = rms ln(max(1 + delta, 0.01)) against the oracle's deposit; forces with
    the multiplier table equal forces without it on delta(k) M(|n|^2)."""
    n, ns, box = 32, 16, 64.0
    g = M.MGPU(n, ns, box, OM)
    pos = lattice_particles(ns, box, amp=1.5, seed=3)
    g.set_particles(pos)
    m = M.MLP(seed=5)
    a, theta = 0.6, 0.5
    # random weights, output layer scaled so that ln M spans about +-0.5
    o = Q.unpack(random_params(31, 0.3))
    xs0 = np.stack([np.linspace(0, 1, 50), np.full(50, a), np.full(50, theta), np.full(50, 0.5)], axis=1)
    o["b4"][:] = 0.0
    y0 = Q.forward(Q.pack(o), xs0)
    o["W4"] *= 0.5 / max(np.abs(y0 - np.median(y0)).max(), 1e-6)
    o["b4"][:] = -np.median(Q.forward(Q.pack(o), xs0))
    p = Q.pack(o).astype(np.float32).astype(np.float64)
    m.params = p
    g.set_learned_multiplier(m, a, theta)
    g.pm_deposit()
    s = g.density_feature()
    d = O.real_view(O.deposit(O.cfg(n, ns, box, OM), pos), n)
    s_ref = np.sqrt(np.mean(np.log(np.maximum(1.0 + d, 0.01)) ** 2))
    assert abs(s - s_ref) <= 1e-5 * s_ref, (s, s_ref)
    g.pm_transform()
    dk = g.get_mesh(M.MESH_DENSITY_K)
    g.forces()
    f1 = [g.get_mesh(q) for q in (M.MESH_FORCE_X, M.MESH_FORCE_Y, M.MESH_FORCE_Z)]
    # the same multiplier from the oracle, per |n|^2 of every stored mode
    f = np.fft.fftfreq(n) * n
    n2 = (f[:, None, None] ** 2 + f[None, :, None] ** 2 + np.arange(n // 2 + 1)[None, None, :] ** 2).astype(int)
    n2max = 3 * (n // 2) ** 2
    xs = np.stack([np.minimum(np.sqrt(np.arange(n2max + 1, dtype=np.float32)) / np.float32(n / 2), 1.0),
                   np.full(n2max + 1, a), np.full(n2max + 1, theta), np.full(n2max + 1, s)], axis=1)
    mult = np.exp(Q.forward(p, xs.astype(np.float32).astype(np.float64)))
    assert mult.max() / mult.min() > 1.2  # a multiplier that does something
    g.set_learned_multiplier(None, a, theta)
    g.set_mesh(M.MESH_DENSITY_K, dk * mult[n2])
    g.forces()
    f2 = [g.get_mesh(q) for q in (M.MESH_FORCE_X, M.MESH_FORCE_Y, M.MESH_FORCE_Z)]
    for u, v in zip(f1, f2):
        assert rel_err(u, v) < 3e-3
    g.set_learned_multiplier(None, -1.0, 0.0)
    g.close()
    m.close()
