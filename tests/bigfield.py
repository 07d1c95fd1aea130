"""Separable test fields whose discrete Fourier transforms are known exactly.

A field t(x, y, z) = sum_r g_r(x) h_r(y) u_r(z) has the unnormalised r2c half
spectrum T(kx, ky, kz) = sum_r G_r(kx) H_r(ky) U_r(kz), with G, H, U the 1D
DFTs of the factors.  The 1D transforms are computed in FP64 by numpy, so the
3D spectrum of a field on a 1024^3 or 1536^3 mesh is known to FP64 accuracy
without any 3D FFT on the host.  Random factors make every mode and every
cell non-zero, so a tile of the 3D transform that is skipped, transposed or
mis-twiddled shows up in the comparison.

The host-side work (building the 3D arrays, comparing them) is split over
x-slabs and run on a thread pool; numpy releases the GIL in those loops.
"""
import os
from concurrent.futures import ThreadPoolExecutor

import numpy as np

WORKERS = max(1, min(16, os.cpu_count() or 1))


def _slabs(n, parts):
    step = max(1, -(-n // parts))
    return [(a, min(n, a + step)) for a in range(0, n, step)]


def pmap(fn, n):
    """fn(a, b) over x-slabs [a, b) of [0, n) on the thread pool; results in order."""
    with ThreadPoolExecutor(WORKERS) as ex:
        return list(ex.map(lambda ab: fn(*ab), _slabs(n, 8 * WORKERS)))


def wavenumbers(n):
    """Signed integer wavenumber of each index along x and y (Nyquist +N/2, as
    Forces does at src/auxPM.c:468-476) and along the half axis z."""
    i = np.arange(n)
    return np.where(i > n // 2, i - n, i).astype(np.float64), np.arange(n // 2 + 1, dtype=np.float64)


class Separable:
    """Random rank-`rank` separable real field on an n^3 mesh and its exact
    r2c half spectrum.  Axes listed in zero_mean_axes get factors with zero
    mean, so the spectrum vanishes on the k_a = 0 plane."""

    def __init__(self, n, rank=2, seed=0, zero_mean_axes=()):
        rng = np.random.default_rng(seed)
        self.n = n
        self.f = [rng.standard_normal((rank, n)) for _ in range(3)]
        for a in zero_mean_axes:
            self.f[a] -= self.f[a].mean(axis=1, keepdims=True)
        g, h, u = self.f
        self.G = np.fft.fft(g, axis=1)
        self.H = np.fft.fft(h, axis=1)
        self.U = np.fft.rfft(u, axis=1)
        self.rank = rank

    def real_slab(self, a, b, scale=1.0):
        g, h, u = self.f
        out = np.zeros((b - a, self.n, self.n))
        for r in range(self.rank):
            out += (scale * g[r, a:b])[:, None, None] * (h[r][:, None] * u[r][None, :])[None]
        return out

    def spectrum_slab(self, a, b):
        out = np.zeros((b - a, self.n, self.n // 2 + 1), np.complex128)
        for r in range(self.rank):
            out += self.G[r, a:b][:, None, None] * (self.H[r][:, None] * self.U[r][None, :])[None]
        return out

    def real(self, scale=1.0):
        out = np.empty((self.n, self.n, self.n))

        def fill(a, b):
            out[a:b] = self.real_slab(a, b, scale)

        pmap(fill, self.n)
        return out

    def spectrum(self):
        out = np.empty((self.n, self.n, self.n // 2 + 1), np.complex128)

        def fill(a, b):
            out[a:b] = self.spectrum_slab(a, b)

        pmap(fill, self.n)
        return out

    def force_source(self, axis, box):
        """The density spectrum S whose Forces component `axis` equals N^3 t(x):
        Forces forms F_a(k) = i n_a (-1/n^2) S / (N^3 2 pi/Box) (src/auxPM.c:468-486),
        so S = T i n^2 N^3 (2 pi/Box) / n_a, zero where n_a = 0 (T vanishes
        there when the axis-`axis` factor has zero mean)."""
        n = self.n
        dxy, dz = wavenumbers(n)
        c = float(n) ** 3 * 2.0 * np.pi / box
        out = np.empty((n, n, n // 2 + 1), np.complex128)

        def fill(a, b):
            t = self.spectrum_slab(a, b)
            n2 = (dxy[a:b] ** 2)[:, None, None] + (dxy ** 2)[None, :, None] + (dz ** 2)[None, None, :]
            na = [dxy[a:b][:, None, None], dxy[None, :, None], dz[None, None, :]][axis]
            na = np.broadcast_to(na, t.shape)
            with np.errstate(divide="ignore", invalid="ignore"):
                s = np.where(na != 0, t * 1j * n2 * c / np.where(na != 0, na, 1.0), 0.0)
            out[a:b] = s

        pmap(fill, n)
        return out


def compare(got, expect_slab, n):
    """max |got - expect| / max |expect| and the rms ratio, slab by slab;
    expect_slab(a, b) gives the expected values of x-slab [a, b)."""

    def one(a, b):
        e = expect_slab(a, b)
        d = got[a:b] - e
        return (float(np.abs(d).max()), float(np.abs(e).max()), float((np.abs(d) ** 2).sum()),
                float((np.abs(e) ** 2).sum()))

    r = pmap(one, n)
    dmax = max(x[0] for x in r)
    emax = max(x[1] for x in r)
    drms = np.sqrt(sum(x[2] for x in r))
    erms = np.sqrt(sum(x[3] for x in r))
    return dmax / emax, drms / erms
