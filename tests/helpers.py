"""Shared helpers of the test-suite: synthetic particle sets and comparisons."""
import numpy as np


def lattice_particles(ns, box, amp, seed=1, clustered=True):
    """Particles displaced from the Ns^3 lattice (float32 SoA [3][Ns^3]).

    A smooth random displacement field plus a few compact clumps gives a
    density field with voids and peaks like an evolved COLA snapshot."""
    rng = np.random.default_rng(seed)
    q = (np.indices((ns, ns, ns)).reshape(3, -1).astype(np.float64) + 0.0) * (box / ns)
    disp = np.zeros_like(q)
    for _ in range(6):
        k = rng.integers(1, 4, size=3)
        ph = rng.uniform(0, 2 * np.pi)
        a = rng.normal(size=3) * amp
        disp += a[:, None] * np.sin(2 * np.pi * (k[:, None] * q).sum(0) / box + ph)[None, :]
    pos = q + disp
    if clustered:
        npart = pos.shape[1]
        for c in range(4):
            sel = rng.choice(npart, npart // 20, replace=False)
            centre = rng.uniform(0, box, size=3)
            pos[:, sel] = centre[:, None] + rng.normal(size=(3, sel.size)) * (box / ns) * 0.7
    pos = np.mod(pos, box)
    pos = pos.astype(np.float32)
    pos[pos >= np.float32(box)] = 0.0
    return np.ascontiguousarray(pos)


def rel_err(a, b):
    """max |a - b| / max |b| (real or complex)."""
    a = np.asarray(a)
    b = np.asarray(b)
    a = a.astype(np.complex128 if np.iscomplexobj(a) else np.float64)
    b = b.astype(np.complex128 if np.iscomplexobj(b) else np.float64)
    return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-300))
