"""numpy restatement of the learned Green's-function multiplier's network (test
infrastructure only: tests/ import it as the checker of the GPU kernels).

The network 4 -> 64 -> 64 -> 64 -> 1 with GELU activations, parameters packed
as W1[64][4], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64], b4
(W[out][in]), loss sum w (y - t)^2 / sum w.  With bf16=True the arithmetic
rounds at the points the GPU kernels round (weights to BF16 for the forward;
the hidden activations entering the two 64x64 products and the pre-activation
gradients entering the backward products to BF16; FP64 elsewhere instead of the
kernels' FP32 accumulation), with the gradient taken straight through the
roundings.  With bf16=False it is the exact FP64 function, whose gradient is
pinned by finite differences.  The reference has no such network: it leaves
the grid_corr factor of Forces at 1 (src/auxPM.c:488-498).
"""
import math

import numpy as np

try:
    from scipy.special import erf as _erf
except ImportError:  # pragma: no cover
    _erf = np.vectorize(math.erf)

NPARAMS = 8705
SIZES = [("W1", (64, 4)), ("b1", (64,)), ("W2", (64, 64)), ("b2", (64,)), ("W3", (64, 64)), ("b3", (64,)),
         ("W4", (64,)), ("b4", (1,))]


def unpack(p):
    out, o = {}, 0
    for name, shape in SIZES:
        n = int(np.prod(shape))
        out[name] = np.asarray(p[o:o + n], dtype=np.float64).reshape(shape)
        o += n
    assert o == NPARAMS
    return out


def pack(d):
    return np.concatenate([np.asarray(d[name], dtype=np.float64).ravel() for name, _ in SIZES])


def bf16(a):
    """Round to BF16 (nearest even) through FP32, as the kernels do."""
    f = np.asarray(a, dtype=np.float32)
    u = f.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def gelu(z):
    return 0.5 * z * (1.0 + _erf(z / math.sqrt(2.0)))


def dgelu(z):
    return 0.5 * (1.0 + _erf(z / math.sqrt(2.0))) + z * np.exp(-0.5 * z * z) / math.sqrt(2.0 * math.pi)


def _id(a):
    return np.asarray(a, dtype=np.float64)


def forward(p, x, bf16_on=True, keep=False):
    r = bf16 if bf16_on else _id
    P = unpack(p)
    W1, W2, W3, W4 = r(P["W1"]), r(P["W2"]), r(P["W3"]), r(P["W4"])
    x = np.asarray(x, dtype=np.float64)
    z1 = x @ W1.T + P["b1"]
    h1 = r(gelu(z1))
    z2 = h1 @ W2.T + P["b2"]
    h2 = r(gelu(z2))
    z3 = h2 @ W3.T + P["b3"]
    g3 = gelu(z3)
    y = g3 @ W4 + P["b4"][0]
    if keep:
        return y, dict(x=x, z1=z1, h1=h1, z2=z2, h2=h2, z3=z3, g3=g3, W2=W2, W3=W3, W4=W4)
    return y


def loss_and_grad(p, x, t, w, bf16_on=True):
    r = bf16 if bf16_on else _id
    y, c = forward(p, x, bf16_on, keep=True)
    t = np.asarray(t, dtype=np.float64)
    w = np.asarray(w, dtype=np.float64)
    W = w.sum()
    e = y - t
    loss = float((w * e * e).sum() / W)
    dy = 2.0 * w * e / W
    g = {}
    g["W4"] = c["g3"].T @ dy
    g["b4"] = np.array([dy.sum()])
    dz3 = dy[:, None] * c["W4"][None, :] * dgelu(c["z3"])
    g["b3"] = dz3.sum(0)
    dz3r = r(dz3)
    g["W3"] = dz3r.T @ c["h2"]
    dz2 = (dz3r @ c["W3"]) * dgelu(c["z2"])
    g["b2"] = dz2.sum(0)
    dz2r = r(dz2)
    g["W2"] = dz2r.T @ c["h1"]
    dz1 = (dz2r @ c["W2"]) * dgelu(c["z1"])
    g["b1"] = dz1.sum(0)
    g["W1"] = dz1.T @ c["x"]
    return loss, pack(g)
