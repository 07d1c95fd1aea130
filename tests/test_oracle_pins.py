"""Pins of the CPU oracle (liboracle.so) against closed forms, brute force and
published known answers.  The oracle is what the GPU path is checked against,
so each test here is chosen so that a plausible slip in the restatement (a
dropped term, a wrong sign or index, a transposed operand, a wrong Nyquist
convention) fails it.  None of these tests need a GPU.

Pinned:
  RANLUX ranlxd1         GSL's own known answer (seed 1, 10000th value)
  3D r2c / c2r           numpy.fft (independent FFT), Parseval, FFTW c2r semantics
  CIC deposit            closed-form weights of one particle, uniform lattice,
                         brute force (numpy scatter), ghost-slab wrap
  CIC gather             adjointness with the deposit
  Green's function       analytic force of plane waves, +N/2 Nyquist convention
  P(k) estimator         plane waves of known amplitude, mode counts by brute force
  MG fifth force         f(R) unscreened multiplier in physical units, screening
                         limits and an analytic screened plane wave, DGP density
                         screening of a plane wave, JBD G_eff scaling
  2LPT                   single plane wave (zero source), two orthogonal waves
                         (analytic source and displacement)
  Kick / Drift           free streaming with zero force, periodic wrap; the COLA
                         identity (zero kick when the PM force is the LPT force,
                         ZA and 2LPT slots); a 1D Zel'dovich plane wave run through
                         deposit, forces, gather, Kick and Drift follows
                         x = q + D(a) Psi(q) before shell crossing
  Seed table             hand-derived draw order of src/2LPT.c:262-271 for N = 2, 4;
                         the table of N is the table of 2N's inner shells
  IC delta(k)            amplitude-fixed modes |delta|^2 = P / Box^3, inverted = -delta,
                         Hermitian kz = 0 plane, zero Nyquist planes and DC, exact cube
                         and sphere cuts, draws consumed by skipped modes (the phases
                         do not depend on Nsample), phase of a mode = 2 pi x the
                         column stream's draw of that mode, |delta|^2 Box^3 / P
                         exponentially distributed
  LPT fields             g = 1 gives the ZA field; the 2LPT field of the stored
                         source with g = -3/7/N^3 gives the 2LPT displacement;
                         plane waves with a k-dependent growth g(|k|)
"""
import numpy as np
import pytest

import oracle as O

INV_H0 = 2997.92458  # Mpc/h


def cos_mesh(n, modes):
    """sum_i amp_i cos(2 pi m_i . j / n + phase_i) on an n^3 grid, in a padded oracle mesh."""
    j = np.indices((n, n, n)).astype(np.float64)
    f = np.zeros((n, n, n))
    for amp, m, ph in modes:
        f += amp * np.cos(2 * np.pi * (m[0] * j[0] + m[1] * j[1] + m[2] * j[2]) / n + ph)
    d = O.new_mesh(n)
    O.real_view(d, n)[...] = f
    return d, f


def theta(n, m):
    j = np.indices((n, n, n)).astype(np.float64)
    return 2 * np.pi * (m[0] * j[0] + m[1] * j[1] + m[2] * j[2]) / n


# ---------------------------------------------------------------- RNG

def test_ranlxd1_gsl_known_answer():
    # GSL's rng test-suite value for gsl_rng_ranlxd1, seed 1, after 10000 draws
    assert O.ranlxd_get_nth(1, 10000) == 1998227290


def test_ranlxd1_uniform_range_and_seed_dependence():
    a = O.ranlxd_stream(5001, 2000)
    b = O.ranlxd_stream(5002, 2000)
    assert a.min() >= 0.0 and a.max() < 1.0
    assert abs(a.mean() - 0.5) < 0.03
    assert not np.array_equal(a, b)
    # 48-bit mantissa: every value is a multiple of 2^-48
    assert np.all(np.floor(a * 2.0 ** 48) == a * 2.0 ** 48)


# ---------------------------------------------------------------- FFT

@pytest.mark.parametrize("n", [2, 4, 6, 8, 10, 12, 16, 18, 20, 24, 30, 32])
def test_r2c_matches_numpy(n):
    rng = np.random.default_rng(n)
    d = O.new_mesh(n)
    x = rng.normal(size=(n, n, n))
    O.real_view(d, n)[...] = x
    O.fft_r2c_3d(n, d)
    ref = np.fft.rfftn(x)
    np.testing.assert_allclose(O.complex_view(d, n), ref, rtol=0, atol=1e-12 * np.abs(ref).max())


@pytest.mark.parametrize("n", [4, 8, 12, 16, 20])
def test_c2r_is_fftw_backward(n):
    """Unnormalised c2r of an arbitrary (non-Hermitian) half spectrum: complex
    backward transforms along x and y, then a c2r along z that uses only the
    real part of the kz = 0 and kz = N/2 entries (what FFTW's rdft2 does)."""
    rng = np.random.default_rng(100 + n)
    h = n // 2 + 1
    c = rng.normal(size=(n, n, h)) + 1j * rng.normal(size=(n, n, h))
    d = O.new_mesh(n)
    O.complex_view(d, n)[...] = c
    O.fft_c2r_3d(n, d)
    ref = np.fft.irfftn(c, s=(n, n, n), axes=(0, 1, 2)) * n ** 3
    np.testing.assert_allclose(O.real_view(d, n), ref, rtol=0, atol=1e-11 * np.abs(ref).max())


@pytest.mark.parametrize("n", [8, 12])
def test_roundtrip_and_parseval(n):
    rng = np.random.default_rng(7)
    x = rng.normal(size=(n, n, n))
    d = O.new_mesh(n)
    O.real_view(d, n)[...] = x
    O.fft_r2c_3d(n, d)
    c = O.complex_view(d, n)
    w = np.full(c.shape[2], 2.0)
    w[0] = 1.0
    if n % 2 == 0:
        w[-1] = 1.0
    assert abs((np.abs(c) ** 2 * w).sum() / n ** 3 - (x ** 2).sum()) < 1e-10 * (x ** 2).sum()
    O.fft_c2r_3d(n, d)
    np.testing.assert_allclose(O.real_view(d, n) / n ** 3, x, atol=1e-12)


# ---------------------------------------------------------------- CIC deposit / gather

def brute_cic(n, ns, box, pos):
    """Independent CIC: numpy scatter-add of the 8 weights with periodic indices."""
    rho = np.zeros((n, n, n))
    wpar = (n / ns) ** 3
    x = pos.astype(np.float64) * (n / box)
    i0 = np.floor(x).astype(np.int64)
    f = x - i0
    for dx in (0, 1):
        for dy in (0, 1):
            for dz in (0, 1):
                w = (np.where(dx, f[0], 1 - f[0]) * np.where(dy, f[1], 1 - f[1]) * np.where(dz, f[2], 1 - f[2]))
                np.add.at(rho, ((i0[0] + dx) % n, (i0[1] + dy) % n, (i0[2] + dz) % n), w * wpar)
    return rho - 1.0


def test_cic_single_particle_closed_form():
    n, box = 8, 8.0
    p = np.array([[2.25], [5.5], [7.75]], dtype=np.float32)  # cell units (box = n)
    c = O.cfg(n, n, box, 0.3)
    d = O.real_view(O.deposit(c, p), n)
    expect = -np.ones((n, n, n))
    wx = {2: 0.75, 3: 0.25}
    wy = {5: 0.5, 6: 0.5}
    wz = {7: 0.25, 0: 0.75}  # wraps through the periodic boundary
    for ix, a in wx.items():
        for iy, b in wy.items():
            for iz, cc in wz.items():
                expect[ix, iy, iz] += a * b * cc
    np.testing.assert_allclose(d, expect, atol=1e-14)


def test_cic_ghost_slab_wraps_x():
    """A particle in the last x cell deposits into slab 0 through the ghost slab."""
    n, box = 8, 8.0
    p = np.array([[7.5], [1.0], [1.0]], dtype=np.float32)
    d = O.real_view(O.deposit(O.cfg(n, n, box, 0.3), p), n)
    assert d[7, 1, 1] == pytest.approx(-0.5)
    assert d[0, 1, 1] == pytest.approx(-0.5)
    assert (d + 1).sum() == pytest.approx(1.0)


def test_cic_uniform_lattice_is_zero():
    n, box = 16, 32.0
    q = (np.indices((n, n, n)).reshape(3, -1) * (box / n)).astype(np.float32)
    d = O.real_view(O.deposit(O.cfg(n, n, box, 0.3), q), n)
    assert np.abs(d).max() < 1e-14


@pytest.mark.parametrize("n,ns", [(16, 16), (16, 8), (12, 6)])
def test_cic_brute_force(n, ns):
    rng = np.random.default_rng(n * 7 + ns)
    box = 50.0
    pos = rng.uniform(0, box, size=(3, ns ** 3)).astype(np.float32)
    d = O.real_view(O.deposit(O.cfg(n, ns, box, 0.3), pos), n)
    ref = brute_cic(n, ns, box, pos)
    np.testing.assert_allclose(d, ref, atol=1e-11)
    assert abs(d.sum()) < 1e-9


@pytest.mark.parametrize("n,ns", [(32, 48), (24, 64), (96, 48)])
def test_cic_brute_force_threaded_path(n, ns):
    """More than 65536 particles take the oracle's threaded deposit (x-slab blocks in
    two phases): clumps across block edges, particles in the last slab and on x = 0."""
    rng = np.random.default_rng(n + ns)
    box = 40.0
    npart = ns ** 3
    pos = rng.uniform(0, box, size=(3, npart))
    k = npart // 3
    pos[:, :k] = rng.normal(box / 2, box / n, size=(3, k))       # a clump over the middle slabs
    pos[0, k:k + 1000] = rng.uniform(box * (1 - 1.0 / n), box, 1000)  # the last slab
    pos[0, k + 1000:k + 2000] = 0.0
    pos = np.mod(pos, box).astype(np.float32)
    pos[pos >= np.float32(box)] = 0.0
    d = O.real_view(O.deposit(O.cfg(n, ns, box, 0.3), pos), n)
    ref = brute_cic(n, ns, box, pos)
    np.testing.assert_allclose(d, ref, atol=1e-9 * max(1.0, np.abs(ref).max()))


def test_gather_is_adjoint_of_deposit():
    n, ns, box = 16, 8, 40.0
    rng = np.random.default_rng(3)
    pos = rng.uniform(0, box, size=(3, ns ** 3)).astype(np.float32)
    c = O.cfg(n, ns, box, 0.3)
    rho = O.real_view(O.deposit(c, pos), n) + 1.0
    f = []
    for a in range(3):
        m = O.new_mesh(n)
        O.real_view(m, n)[...] = rng.normal(size=(n, n, n))
        m[n * n * 2 * (n // 2 + 1):] = m[: n * 2 * (n // 2 + 1)]  # ghost slab = slab 0
        f.append(m)
    disp, s = O.gather(c, pos, f)
    wpar = (n / ns) ** 3
    for a in range(3):
        lhs = disp[a].astype(np.float64).sum()
        rhs = (O.real_view(f[a], n) * rho).sum() / wpar
        assert abs(lhs - rhs) < 1e-5 * np.abs(O.real_view(f[a], n)).sum()
        assert s[a] == pytest.approx(disp[a].astype(np.float64).mean(), rel=1e-6)


# ---------------------------------------------------------------- Green's function / forces

@pytest.mark.parametrize("m", [(1, 0, 0), (0, 2, 0), (0, 0, 3), (1, -2, 3), (-3, 1, 1), (2, 5, -1)])
def test_force_of_plane_wave(m):
    """delta = eps cos(k.x): F_a = eps (m_a / |m|^2) (Box / 2 pi) sin(k.x)."""
    n, box, eps = 16, 100.0, 0.3
    d, _ = cos_mesh(n, [(eps, m, 0.0)])
    c = O.cfg(n, n, box, 0.3)
    O.fft_r2c_3d(n, d)
    f = O.forces(c, d, None)
    m2 = float(np.dot(m, m))
    s = np.sin(theta(n, m))
    for a in range(3):
        np.testing.assert_allclose(O.real_view(f[a], n), eps * m[a] / m2 * box / (2 * np.pi) * s, atol=1e-12)


def test_force_nyquist_is_plus_half_n():
    """A wave at the x-Nyquist index with kz > 0: the surviving force term uses n_x = +N/2."""
    n, box, eps = 16, 64.0, 0.2
    m = (n // 2, 0, 3)
    d, _ = cos_mesh(n, [(eps, m, 0.0)])
    O.fft_r2c_3d(n, d)
    f = O.forces(O.cfg(n, n, box, 0.3), d, None)
    m2 = float(np.dot(m, m))
    s = np.sin(theta(n, m))
    np.testing.assert_allclose(O.real_view(f[0], n), eps * (n / 2) / m2 * box / (2 * np.pi) * s, atol=1e-12)
    np.testing.assert_allclose(O.real_view(f[2], n), eps * 3 / m2 * box / (2 * np.pi) * s, atol=1e-12)


def test_forces_with_phi_add_linearly():
    n, box = 12, 60.0
    d, _ = cos_mesh(n, [(0.1, (1, 2, 0), 0.3), (0.05, (0, 1, 4), 1.1)])
    O.fft_r2c_3d(n, d)
    phi, _ = cos_mesh(n, [(0.07, (2, 0, 1), 0.0)])
    O.fft_r2c_3d(n, phi)
    c = O.cfg(n, n, box, 0.3)
    both = O.forces(c, d.copy(), phi)
    s = d.copy()
    s[: n * n * (n // 2 + 1) * 2] += phi[: n * n * (n // 2 + 1) * 2]
    ref = O.forces(c, s, None)
    for a in range(3):
        np.testing.assert_allclose(O.real_view(both[a], n), O.real_view(ref[a], n), atol=1e-13)


# ---------------------------------------------------------------- P(k)

def mode_counts(n, nbins, kmin, kmax):
    f = np.fft.fftfreq(n) * n
    kk = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + np.abs(f)[None, None, :] ** 2)
    b = np.floor((kk - kmin) / (kmax - kmin) * nbins + 0.5).astype(int)
    b = b[(b >= 0) & (b < nbins)]
    return np.bincount(b, minlength=nbins).astype(float)


@pytest.mark.parametrize("m", [(3, 0, 0), (0, 3, 0), (0, 0, 3), (2, -1, 2)])
def test_pofk_plane_wave(m):
    """eps cos(k.x) puts eps^2/2 Box^3 / W(k) / nmodes(bin) into the bin of |m|,
    with W the CIC window prod sinc^4(pi m_a / N); all other bins hold zero."""
    n, ns, box, eps, nbins = 16, 8, 200.0, 0.1, 16
    d, _ = cos_mesh(n, [(eps, m, 0.4)])
    O.fft_r2c_3d(n, d)
    c = O.cfg(n, ns, box, 0.3)
    p, k, nm = O.pofk(c, d, nbins, 0, 0.0, float(n), 0)
    counts = mode_counts(n, nbins, 0.0, float(n))
    np.testing.assert_array_equal(nm, counts)
    mm = np.sqrt(np.dot(m, m))
    b = int(np.floor(mm / n * nbins + 0.5))
    w = np.prod([np.sinc(mi / n) ** 4 for mi in m])
    expect = np.zeros(nbins)
    expect[b] = eps ** 2 / 2 * box ** 3 / w / counts[b]
    np.testing.assert_allclose(p, expect, rtol=1e-10, atol=1e-14 * expect.max())
    # shot noise is subtracted from every populated bin
    p2, _, _ = O.pofk(c, d, nbins, 0, 0.0, float(n), 1)
    np.testing.assert_allclose(p2[counts > 0], (expect - (box / ns) ** 3)[counts > 0], rtol=1e-12)


def test_pofk_kmean_is_mode_average():
    n, box, nbins = 12, 30.0, 6
    d = O.new_mesh(n)
    O.fft_r2c_3d(n, d)
    _, k, nm = O.pofk(O.cfg(n, n, box, 0.3), d, nbins, 0, 0.0, float(n), 0)
    f = np.fft.fftfreq(n) * n
    kk = np.sqrt(f[:, None, None] ** 2 + f[None, :, None] ** 2 + np.abs(f)[None, None, :] ** 2).ravel()
    b = np.floor(kk / n * nbins + 0.5).astype(int)
    for i in range(nbins):
        sel = b == i
        if sel.any():
            assert k[i] == pytest.approx(kk[sel].mean() * 2 * np.pi / box, rel=1e-12)


def test_adjust_pofk_defaults():
    assert O.adjust_pofk(64, 0, 0, 0, 0.0, 0.0) == (64, 0, 0, 0.0, 64.0)
    assert O.adjust_pofk(64, 10, 1, 1, 0.0, 0.0) == (10, 1, 1, 1.0, 64.0)
    assert O.adjust_pofk(64, 10, 7, 5, -1.0, 500.0) == (10, 0, 1, 0.0, 64.0)
    assert O.adjust_pofk(64, 10, 0, 0, 2.0, 30.0) == (10, 0, 0, 2.0, 30.0)


# ---------------------------------------------------------------- RSD multipoles

@pytest.mark.parametrize("m,mu2", [((0, 0, 3), 1.0), ((3, 0, 0), 0.0), ((0, 2, 0), 0.0), ((2, 0, 2), 0.5),
                                   ((0, -2, 2), 0.5)])
def test_rsd_multipoles_plane_wave(m, mu2):
    """A plane wave with mu^2 = m_z^2/|m|^2 (line of sight = mesh z) gives
    P2 = 5/2 (3 mu^2 - 1) P0 and P4 = 9/8 (35 mu^4 - 30 mu^2 + 3) P0 in its bin,
    with P0 the plain P(k) of the same field; zero elsewhere."""
    n, ns, box, eps, nbins = 16, 8, 200.0, 0.1, 16
    d, _ = cos_mesh(n, [(eps, m, 0.4)])
    O.fft_r2c_3d(n, d)
    c = O.cfg(n, ns, box, 0.3)
    p0, p2, p4, k, nm = O.rsd_multipoles(c, d, nbins, 0, 0.0, float(n), 0)
    pk, _, nm_pk = O.pofk(c, d, nbins, 0, 0.0, float(n), 0)
    np.testing.assert_array_equal(nm, nm_pk)
    np.testing.assert_allclose(p0, pk, rtol=1e-12, atol=1e-14 * pk.max())
    b = int(np.floor(np.sqrt(np.dot(m, m)) / n * nbins + 0.5))
    np.testing.assert_allclose(p2[b], 2.5 * (3 * mu2 - 1) * p0[b], rtol=1e-10)
    np.testing.assert_allclose(p4[b], 9.0 / 8.0 * (35 * mu2 ** 2 - 30 * mu2 + 3) * p0[b], rtol=1e-10)
    other = np.arange(nbins) != b
    assert np.abs(p2[other]).max() < 1e-12 * abs(p0[b]) and np.abs(p4[other]).max() < 1e-12 * abs(p0[b])
    # k of a bin is its nominal value (k_from_index), not the mode average
    np.testing.assert_allclose(k[nm > 0], (np.arange(nbins) * n / nbins * 2 * np.pi / box)[nm > 0], rtol=1e-14)
    # shot noise comes off all three multipoles
    q0, q2, q4, _, _ = O.rsd_multipoles(c, d, nbins, 0, 0.0, float(n), 1)
    sn = (box / ns) ** 3
    for a, b_ in ((q0, p0), (q2, p2), (q4, p4)):
        np.testing.assert_allclose(a[nm > 0], (b_ - sn)[nm > 0], rtol=1e-12, atol=1e-12 * sn)


def rsd_particles(ns, box, seed):
    rng = np.random.default_rng(seed)
    pos = rng.uniform(0, box, size=(3, ns ** 3)).astype(np.float32)
    z = np.zeros((3, ns ** 3), np.float32)
    return pos, z


def test_rsd_deposit_axis_mapping_and_shifts():
    """Zero velocity: the z line of sight is the plain deposit and the y line
    of sight the plain deposit of (x, z, y).  A velocity of an integer number
    of cells rolls the field along the line of sight, for each form of V:
    plain Vel, Vel + (dD/dy summed over orders) and Vel + D dDdy + D2 dD2dy."""
    n, ns, box = 16, 8, 32.0
    c = O.cfg(n, ns, box, 0.3)
    pos, z = rsd_particles(ns, box, 7)
    ref = O.real_view(O.deposit(c, pos), n)
    dz = O.real_view(O.deposit_rsd(c, pos, z, z, z, 2, 1.0, 0, 0), n)
    np.testing.assert_array_equal(dz, ref)
    swapped = np.ascontiguousarray(pos[[0, 2, 1]])
    dy = O.real_view(O.deposit_rsd(c, pos, z, z, z, 1, 1.0, 0, 0), n)
    np.testing.assert_array_equal(dy, O.real_view(O.deposit(c, swapped), n))
    one = np.ones_like(pos)
    # 2 cells along z: vnorm = 1 cell per unit, V = 2 in three ways
    cases = [dict(vel=2 * one, l0=z, l1=z, sd=0, cola=0, s1=0.0, s2=0.0),
             dict(vel=one, l0=z, l1=one, sd=1, cola=1, s1=0.0, s2=0.0),
             dict(vel=one, l0=one, l1=2 * one, sd=0, cola=1, s1=0.5, s2=0.25)]
    for cs in cases:
        d2 = O.real_view(O.deposit_rsd(c, pos, cs["vel"], cs["l0"], cs["l1"], 2, 1.0, cs["sd"], cs["cola"], cs["s1"],
                                       cs["s2"]), n)
        np.testing.assert_array_equal(d2, np.roll(ref, 2, axis=2)), cs
    # and along y for the y line of sight (mesh z after the swap)
    d3 = O.real_view(O.deposit_rsd(c, pos, 2 * one, z, z, 1, 1.0, 0, 0), n)
    np.testing.assert_array_equal(d3, np.roll(O.real_view(O.deposit(c, swapped), n), 2, axis=2))


# ---------------------------------------------------------------- modified gravity

def mgd(**kw):
    d = dict(a=1.0, coupling=0.0, mass2=0.0, phi_crit=0.0, dgp_fac=0.0, rsmooth=1.0, geff=1.0)
    d.update(kw)
    return O.MG(**d)


def newton_potential(n, box, om, a, eps, m):
    """Phi_N(x) in the units of the screening test: nabla^2 Phi = 1.5 Om / a * delta with k in units of H0."""
    kphys = 2 * np.pi / box * np.sqrt(np.dot(m, m)) * INV_H0
    return -1.5 * om / a * eps * np.cos(theta(n, m)) / kphys ** 2


def test_fofr_unscreened_multiplier_physical_units():
    """Force ratio (1 + 2 beta^2 k^2 / (k^2 + a^2 m^2)) with k in h/Mpc and m^2 in H0^2 units."""
    n, box, eps, om = 16, 256.0, 0.05, 0.3
    a, mass2, coup = 0.6, 900.0, 1.0 / 3.0
    for m in [(1, 0, 0), (0, 3, 2), (5, 1, 0)]:
        c = O.cfg(n, n, box, om, model=O.FOFR, mg_active=1, include_screening=0)
        d, dx = cos_mesh(n, [(eps, m, 0.0)])
        dxfull = d.copy()
        O.fft_r2c_3d(n, d)
        phik = O.fifth_force(c, mgd(a=a, coupling=coup, mass2=mass2), d, dxfull)
        f = O.forces(c, d, phik)
        k2 = (2 * np.pi / box) ** 2 * np.dot(m, m)
        ratio = 1 + coup * k2 / (k2 + a * a * mass2 / INV_H0 ** 2)
        m2 = float(np.dot(m, m))
        s = np.sin(theta(n, m))
        for ax in range(3):
            np.testing.assert_allclose(O.real_view(f[ax], n), ratio * eps * m[ax] / m2 * box / (2 * np.pi) * s,
                                       atol=1e-12)


@pytest.mark.parametrize("phi_crit_scale", [1e30, 0.3])
def test_fofr_potential_screening_plane_wave(phi_crit_scale):
    """delta_eff = delta * min(1, |phi_crit / Phi_N|) where Phi_N < 0, else delta;
    phi(k) = 2 beta^2 delta_eff(k) n^2 / (n^2 + a^2 m^2 (Box / 2 pi c/H0)^2)."""
    n, box, eps, om = 16, 200.0, 0.8, 0.3
    a, mass2, coup = 0.5, 2.0e4, 1.0 / 3.0
    m = (1, 2, 0)
    phin = newton_potential(n, box, om, a, eps, m)
    pc = phi_crit_scale * np.abs(phin).max()
    c = O.cfg(n, n, box, om, model=O.FOFR, mg_active=1, include_screening=1)
    d, dx = cos_mesh(n, [(eps, m, 0.0)])
    dxfull = d.copy()
    O.fft_r2c_3d(n, d)
    phik = O.fifth_force(c, mgd(a=a, coupling=coup, mass2=mass2, phi_crit=pc), d, dxfull)
    s = np.where(phin < 0, np.minimum(1.0, np.abs(pc / np.where(phin == 0, 1.0, phin))), 1.0)
    eff = np.fft.rfftn(dx * s)
    f = np.fft.fftfreq(n) * n
    n2 = f[:, None, None] ** 2 + f[None, :, None] ** 2 + (np.arange(n // 2 + 1) ** 2)[None, None, :]
    mt2 = a * a * mass2 / (2 * np.pi * INV_H0 / box) ** 2
    expect = coup * eff * n2 / (n2 + mt2)
    expect[0, 0, 0] = 0
    got = O.complex_view(phik, n)
    assert np.abs(got - expect).max() < 1e-10 * np.abs(expect).max()
    if phi_crit_scale < 1:
        assert s.min() < 0.5  # the test screens


def test_dgp_density_screening_plane_wave():
    """delta_R = Gaussian-smoothed delta; delta_eff = delta * coupling * 2 (sqrt(1+f) - 1) / f, f = fac (1 + delta_R)."""
    n, box, eps, om = 16, 100.0, 0.6, 0.3
    a, coup, fac, R = 0.7, 0.12, 0.3, 3.0
    m = (2, 0, 1)
    c = O.cfg(n, n, box, om, model=O.DGP, mg_active=1, include_screening=1, filter=O.FILTER_GAUSSIAN)
    d, dx = cos_mesh(n, [(eps, m, 0.0)])
    dxfull = d.copy()
    O.fft_r2c_3d(n, d)
    phik = O.fifth_force(c, mgd(a=a, coupling=coup, dgp_fac=fac, rsmooth=R), d, dxfull)
    k = 2 * np.pi / box * np.sqrt(np.dot(m, m))
    dR = eps * np.exp(-0.5 * (k * R) ** 2) * np.cos(theta(n, m))
    ff = fac * (1 + dR)
    scr = np.where(ff < 1e-5, 1.0, 2 * (np.sqrt(1 + ff) - 1) / ff)
    expect = np.fft.rfftn(dx * coup * scr)
    assert np.abs(O.complex_view(phik, n) - expect).max() < 1e-10 * np.abs(expect).max()


def test_dgp_unscreened_and_jbd_scale_delta():
    n, box, om = 12, 80.0, 0.3
    d, _ = cos_mesh(n, [(0.2, (1, 1, 2), 0.5)])
    O.fft_r2c_3d(n, d)
    c = O.cfg(n, n, box, om, model=O.DGP, mg_active=1, include_screening=0)
    phik = O.fifth_force(c, mgd(coupling=0.2), d.copy(), d.copy())
    np.testing.assert_allclose(O.complex_view(phik, n), 0.2 * O.complex_view(d, n), rtol=1e-15)
    cj = O.cfg(n, n, box, om, model=O.BRANSDICKE, mg_active=1)
    dj = d.copy()
    assert O.fifth_force(cj, mgd(geff=1.07), dj, d.copy()) is None
    np.testing.assert_allclose(O.complex_view(dj, n), 1.07 * O.complex_view(d, n), rtol=1e-15)


# ---------------------------------------------------------------- 2LPT

def wave_spectrum(n, waves):
    """Half spectrum cd with delta(x) = sum amp cos(2 pi m . q / n) under the unnormalised c2r."""
    cd = O.new_mesh(n)
    cv = O.complex_view(cd, n)
    for amp, m in waves:
        i, j, k = (mi % n for mi in m)
        if k == 0:
            cv[i, j, 0] += amp / 2
            cv[(-m[0]) % n, (-m[1]) % n, 0] += amp / 2
        else:
            assert 0 < k < n // 2
            cv[i, j, k] += amp / 2
    return cd


def test_2lpt_single_wave_has_no_source():
    n, box, A = 16, 100.0, 0.05
    m = (0, 0, 2)
    cd = wave_spectrum(n, [(A, m)])
    za, lp = O.ic_displacements(O.cfg(n, n, box, 0.3), cd, False)
    q = 2 * np.pi / box * np.array(m, dtype=float)
    th = theta(n, m).reshape(-1)
    for a in range(3):
        np.testing.assert_allclose(za[a], -A * q[a] / q.dot(q) * np.sin(th), atol=1e-7)
    assert np.abs(lp).max() < 1e-9


def test_2lpt_two_orthogonal_waves():
    """phi1 with nabla^2 phi1 = delta; S = sum_{i<j} (phi1_ii phi1_jj - phi1_ij^2) = A B cos(t1) cos(t2);
    nabla^2 phi2 = S; ZA = -grad phi1, 2LPT field = -(3/7) grad phi2 (mean removed)."""
    n, box, A, B = 16, 100.0, 0.08, 0.05
    m1, m2 = (0, 0, 2), (0, 3, 0)
    cd = wave_spectrum(n, [(A, m1), (B, m2)])
    za, lp = O.ic_displacements(O.cfg(n, n, box, 0.3), cd, False)
    kf = 2 * np.pi / box
    q1, q2 = kf * np.array(m1, float), kf * np.array(m2, float)
    t1, t2 = theta(n, m1).reshape(-1), theta(n, m2).reshape(-1)
    pp, pm = q1 + q2, q1 - q2
    for a in range(3):
        ez = -A * q1[a] / q1.dot(q1) * np.sin(t1) - B * q2[a] / q2.dot(q2) * np.sin(t2)
        np.testing.assert_allclose(za[a], ez, atol=1e-7)
        grad_phi2 = 0.5 * A * B * (pp[a] / pp.dot(pp) * np.sin(t1 + t2) + pm[a] / pm.dot(pm) * np.sin(t1 - t2))
        np.testing.assert_allclose(lp[a], -3.0 / 7.0 * grad_phi2, atol=1e-8)
    assert np.abs(lp).max() > 1e-6


def test_2lpt_sd_source_spectrum_sign():
    """With scale-dependent growth the source is stored as cdelta2 = -S(k)."""
    n, box, A, B = 12, 60.0, 0.08, 0.05
    m1, m2 = (0, 0, 2), (0, 3, 0)
    cd = wave_spectrum(n, [(A, m1), (B, m2)])
    (cd2,) = O.ic_displacements(O.cfg(n, n, box, 0.3), cd, True)
    t1, t2 = theta(n, m1), theta(n, m2)
    S = A * B * np.cos(t1) * np.cos(t2)
    expect = -np.fft.rfftn(S)
    assert np.abs(O.complex_view(cd2, n) - expect).max() < 1e-12 * np.abs(expect).max()


# ---------------------------------------------------------------- Kick / Drift

def test_free_streaming_and_wrap():
    npart, box = 1000, 50.0
    rng = np.random.default_rng(1)
    pos = rng.uniform(0, box, size=(3, npart)).astype(np.float32)
    vel = rng.normal(size=(3, npart)).astype(np.float32) * 5
    disp = np.zeros_like(pos)
    v0 = vel.copy()
    sV = O.kick(vel, disp, None, None, 0.0, 0.0, 0, 0.3, 0.5, 0.5, 0, np.zeros(3))
    np.testing.assert_array_equal(vel, v0)
    np.testing.assert_allclose(sV, v0.astype(np.float64).mean(axis=1), rtol=1e-6)
    p = pos.copy()
    O.drift(p, vel, None, None, 0.0, 0.0, 0, box, 0.7, 0, np.zeros(3))
    expect = np.mod(pos.astype(np.float64) + 0.7 * v0.astype(np.float64), box)
    dd = np.abs(p - expect)
    assert np.minimum(dd, box - dd).max() < 1e-4
    assert p.min() >= 0 and p.max() < box


# ---------------------------------------------------------------- learned multiplier network

def test_mlp_oracle_gradient_is_the_derivative():
    """The exact (FP64, no rounding) loss gradient of oracle/mlp.py against central
    finite differences on 60 random parameters, and the closed form of a network
    whose hidden weights vanish: y = W4 . gelu(b3) + b4 for every input."""
    import mlp as Q
    rng = np.random.default_rng(0)
    p = rng.normal(size=Q.NPARAMS) * 0.3
    x = rng.normal(size=(29, 4))
    t = rng.normal(size=29)
    w = rng.uniform(0.5, 2.0, 29)
    _, g = Q.loss_and_grad(p, x, t, w, bf16_on=False)
    for i in rng.choice(Q.NPARAMS, 60, replace=False):
        h = 1e-6
        up, dn = p.copy(), p.copy()
        up[i] += h
        dn[i] -= h
        fd = (Q.loss_and_grad(up, x, t, w, False)[0] - Q.loss_and_grad(dn, x, t, w, False)[0]) / (2 * h)
        assert abs(fd - g[i]) <= 1e-7 * np.abs(g).max(), i
    d = Q.unpack(p)
    d["W3"][:] = 0.0
    y = Q.forward(Q.pack(d), x, bf16_on=False)
    np.testing.assert_allclose(y, d["W4"] @ Q.gelu(d["b3"]) + d["b4"][0], rtol=1e-12)


def test_bf16_rounding_is_nearest_even():
    import mlp as Q
    one = 1.0
    ulp = 2.0 ** -7  # BF16 has 8 significant bits
    assert Q.bf16(one) == one
    assert Q.bf16(one + ulp / 2) == one                # tie, even neighbour below
    assert Q.bf16(one + 3 * ulp / 2) == one + 2 * ulp  # tie, even neighbour above
    assert Q.bf16(one + ulp / 2 + 2.0 ** -20) == one + ulp
    assert Q.bf16(-3.0) == -3.0
    assert Q.gelu(np.array([0.0]))[0] == 0.0
    np.testing.assert_allclose(Q.dgelu(np.array([0.0])), [0.5])


# ---------------------------------------------------------------- COLA identity and Zel'dovich run

def test_kick_cola_identity_zero_residual():
    """COLA integrates the residual to LPT: when the gathered PM force equals
    the LPT force, -1.5 Omega (Disp - <Disp>) cancels the LPT term (D ddD1 +
    D2 ddD2)/A with ddD1 = 1.5 Omega A D1 and ddD2 = 1.5 Omega A (D2 - D1^2)
    (the growth equations, src/cosmo.c:206-239), so the kick is zero
    (src/main.c:707-716).  Disp here is the force of the 2LPT trajectory in the
    Forces sign convention: -(D1 Psi1 + (D2 - D1^2) Psi2)."""
    rng = np.random.default_rng(4)
    npart, om, A, D1, D2 = 4096, 0.3, 0.4, 0.43, -0.09
    psi1 = rng.normal(size=(3, npart)).astype(np.float32)
    psi2 = (0.1 * rng.normal(size=(3, npart))).astype(np.float32)
    psi1 -= psi1.mean(axis=1, keepdims=True)
    psi2 -= psi2.mean(axis=1, keepdims=True)
    disp = (-(D1 * psi1.astype(np.float64) + (D2 - D1 ** 2) * psi2.astype(np.float64))).astype(np.float32)
    sumD = disp.astype(np.float64).mean(axis=1)
    s1, s2 = 1.5 * om * A * D1, 1.5 * om * A * (D2 - D1 ** 2)
    vel = np.zeros((3, npart), np.float32)
    O.kick(vel, disp.copy(), psi1, psi2, s1, s2, 0, om, 0.7, A, 1, sumD)
    assert np.abs(vel).max() < 1e-6 * 0.7 * 1.5 * om * np.abs(disp).max()
    # and the two terms are each non-zero: without COLA the same force kicks
    vel0 = np.zeros_like(vel)
    O.kick(vel0, disp.copy(), psi1, psi2, s1, s2, 0, om, 0.7, A, 0, sumD)
    assert np.abs(vel0).max() > 0.1 * 0.7 * 1.5 * om * np.abs(disp).max()
    # scale-dependent layout: slot D holds the pre-multiplied ddD field (sum of both orders)
    lpt = (psi1.astype(np.float64) * s1 + psi2.astype(np.float64) * s2).astype(np.float32)
    vel_sd = np.zeros_like(vel)
    O.kick(vel_sd, disp.copy(), lpt, np.zeros_like(lpt), 0.0, 0.0, 1, om, 0.7, A, 1, sumD)
    assert np.abs(vel_sd).max() < 1e-5 * 0.7 * 1.5 * om * np.abs(disp).max()


def test_drift_cola_moves_with_lpt_growth():
    """With zero residual velocity, Drift adds exactly dD1 Psi1 + dD2 Psi2
    (src/main.c:762-769): the particle stays on x = q + D1 Psi1 + D2 Psi2."""
    rng = np.random.default_rng(5)
    npart, box = 2048, 100.0
    q = rng.uniform(10, 90, size=(3, npart))
    psi1 = rng.normal(size=(3, npart)).astype(np.float32)
    psi2 = (0.1 * rng.normal(size=(3, npart))).astype(np.float32)
    D1a, D1b, D2a, D2b = 0.3, 0.42, -0.04, -0.08
    pos = (q + D1a * psi1 + D2a * psi2).astype(np.float32)
    vel = np.zeros((3, npart), np.float32)
    O.drift(pos, vel, psi1, psi2, D1b - D1a, D2b - D2a, 0, box, 0.6, 1, np.zeros(3))
    expect = q + D1b * psi1.astype(np.float64) + D2b * psi2.astype(np.float64)
    assert np.abs(pos - expect).max() < 2e-5
    # scale-dependent layout: the Delta D fields of the two orders, stored pre-multiplied
    pos2 = (q + D1a * psi1 + D2a * psi2).astype(np.float32)
    f1 = ((D1b - D1a) * psi1.astype(np.float64)).astype(np.float32)
    f2 = ((D2b - D2a) * psi2.astype(np.float64)).astype(np.float32)
    O.drift(pos2, vel, f1, f2, 0.0, 0.0, 1, box, 0.6, 1, np.zeros(3))
    assert np.abs(pos2 - expect).max() < 2e-5


def zeldovich_1d(ns, box, eps):
    """Lattice q at cell centres and the 1D displacement Psi_x = eps sin(2 pi q_x / Box)
    (zero in y, z).  With particles off the mesh nodes and moved by less than
    half a cell, the CIC weights are linear in the displacement."""
    q = (np.indices((ns, ns, ns)).reshape(3, -1).astype(np.float64) + 0.5) * (box / ns)
    psi = np.zeros_like(q)
    psi[0] = eps * np.sin(2 * np.pi * q[0] / box)
    return q, psi


def test_pm_force_of_zeldovich_wave_is_lpt_force():
    """In 1D before shell crossing the force on a particle is fixed by the mass
    ordering: Disp = -D Psi(q) for x = q + D Psi.  The PM force (CIC deposit,
    Green's function, CIC gather) reproduces it up to the discreteness of the
    deposit and the gather, under 1 % for one wave on 32 cells."""
    n, ns, box, eps, D = 32, 32, 64.0, 0.05, 0.8  # density contrast 0.4 %: linear
    q, psi = zeldovich_1d(ns, box, eps)
    pos = np.mod(q + D * psi, box).astype(np.float32)
    c = O.cfg(n, ns, box, 0.3)
    d = O.deposit(c, pos)
    O.fft_r2c_3d(n, d)
    disp, _ = O.gather(c, pos, O.forces(c, d, None))
    expect = -D * psi[0]
    np.testing.assert_allclose(disp[0], expect, atol=1e-2 * np.abs(expect).max())
    assert np.abs(disp[0] - expect).max() > 1e-5 * np.abs(expect).max()  # a mesh, not the continuum
    assert np.abs(disp[1:]).max() < 1e-9


def zeldovich_run(nsteps, frac, use_cola):
    """Kick/Drift run of a 1D Zel'dovich wave (amplitude frac Box/2pi, so the
    density contrast at a = 1 is frac) on 32^3 particles and a 32^3 mesh in a
    64 Mpc/h LCDM box from a = 0.05 to 1, with the reference's schedule and the
    host's growth factors and time integrals; returns the displacement from
    x = q + D1(a) Psi(q) and the residual velocity, both over D1 Psi's amplitude."""
    import mgpicola as M
    import workloads as W
    import tempfile
    import os
    n, ns, box = 32, 32, 64.0
    out = tempfile.mkdtemp()
    model, vals, _ = W.values_for("lcdm_64_128", out, Nmesh=n, Nsample=ns, Box=box, Init_Redshift=19.0)
    with open(vals["OutputRedshiftFile"], "w") as f:
        f.write(f"0, {nsteps}\n")
    pf = os.path.join(out, "p.txt")
    M.write_paramfile(pf, vals, model=model)
    hp = M.read_params(pf, model=model)
    cosmo = M.Cosmo(hp)
    eps = frac * box / (2 * np.pi)
    q, psi = zeldovich_1d(ns, box, eps)
    psi32 = psi.astype(np.float32)
    zero = np.zeros_like(psi32)
    c = O.cfg(n, ns, box, hp.Omega)
    A = 1.0 / 20.0
    pos = np.mod(q + cosmo.growth(M.G_D, A) * psi, box).astype(np.float32)
    # COLA: zero residual velocity; otherwise the linear growing-mode velocity dD/dy Psi
    vel = np.zeros_like(pos) if use_cola else (cosmo.growth(M.G_dDdy, A) * psi).astype(np.float32)
    AI = A
    da = (1.0 - A) / nsteps
    for i in range(nsteps):
        AF, AFF = A + 0.5 * da, A + da
        d = O.deposit(c, pos)
        O.fft_r2c_3d(n, d)
        disp, sD = O.gather(c, pos, O.forces(c, d, None))
        sV = O.kick(vel, disp, psi32, zero, cosmo.growth(M.G_ddDddy, A), cosmo.growth(M.G_ddD2ddy, A), 0,
                    hp.Omega, cosmo.Sphi(AI, AF, A), A, use_cola, sD)
        O.drift(pos, vel, psi32, zero, cosmo.growth(M.G_D, AFF) - cosmo.growth(M.G_D, A), 0.0, 0, box,
                cosmo.Sq(A, AFF, AF), use_cola, sV)
        A, AI = AFF, AF
    assert abs(A - 1.0) < 1e-12
    D1 = cosmo.growth(M.G_D, 1.0)
    dx = pos.astype(np.float64) - (q + D1 * psi)
    dx -= box * np.round(dx / box)
    assert np.abs(dx[1:]).max() < 1e-5
    vres = vel if use_cola else vel - cosmo.growth(M.G_dDdy, 1.0) * psi
    return np.abs(dx).max() / (D1 * eps), np.abs(vres).max() / (cosmo.growth(M.G_dDdy, 1.0) * eps)


def test_zeldovich_cola_run_follows_growth():
    """In 1D the Zel'dovich trajectory x = q + D1(a) Psi(q) is exact until
    shell crossing, so a COLA run (src/main.c:394-611, Kick and Drift
    src/main.c:688-787) keeps a residual displacement and velocity at the level
    of the PM force's own error (the CIC window, 0.6 %, and its second-order
    discreteness): 12 steps from a = 0.05 stay within 3 %.  It pins the force,
    COLA, Sphi and Sq terms together.  The same 12 steps without COLA are off by
    a third: the time stepping error that COLA removes."""
    ex, ev = zeldovich_run(12, 0.05, 1)
    print("COLA: position", ex, "velocity", ev)
    assert ex < 0.03 and ev < 0.05
    sx, _ = zeldovich_run(12, 0.05, 0)
    print("standard PM: position", sx)
    assert sx > 0.15


# ---------------------------------------------------------------- seed table and IC delta(k)

def _draws(seed, count):
    return (0x7FFFFFFF * O.ranlxd_stream(seed, count)).astype(np.uint32)


def test_seedtable_draw_order_by_hand():
    """src/2LPT.c:262-271 worked by hand: N = 2 fills (0,0), (1,0), (0,1), (1,1) with
    draws 0, 1, 2, 3 in that order; N = 4 as listed (index i*N + j -> draw)."""
    d = _draws(77, 16)
    assert list(O.seedtable(2, 77)) == [d[0], d[2], d[1], d[3]]
    order4 = [0, 5, 11, 2, 4, 6, 12, 10, 7, 9, 15, 13, 1, 8, 14, 3]
    assert list(O.seedtable(4, 77)) == [d[k] for k in order4]


@pytest.mark.parametrize("n", [6, 8, 12])
def test_seedtable_inner_shells_do_not_depend_on_n(n):
    """Shells are filled from the centre (index 0 and N-1) outwards, so the
    seeds of N's shells reappear at the same places in the table of 2N, and
    every entry gets exactly one draw."""
    a = O.seedtable(n, 5001).reshape(n, n)
    b = O.seedtable(2 * n, 5001).reshape(2 * n, 2 * n)
    h = n // 2
    for i in range(n):
        for j in range(n):
            bi = i if i < h else 2 * n - (n - i)
            bj = j if j < h else 2 * n - (n - j)
            assert a[i, j] == b[bi, bj], (i, j)
    assert sorted(a.ravel()) == sorted(_draws(5001, n * n))


def ic_pk(n, box):
    kf = 2 * np.pi / box
    n2 = np.arange(3 * (n // 2) ** 2 + 1)
    k = kf * np.sqrt(n2)
    return np.where(n2 > 0, 3e3 * k / (1 + (k / 0.1) ** 2.5), 0.0)


def signed(n):
    i = np.arange(n)
    return np.where(i < n // 2, i, i - n)


@pytest.mark.parametrize("n,ns,sphere", [(16, 16, 0), (16, 8, 0), (16, 8, 1), (12, 6, 1)])
def test_ic_delta_fixed_amplitude_cut_and_symmetry(n, ns, sphere):
    box = 3.0 * n
    pk = ic_pk(n, box)
    c = O.cfg(n, ns, box, 0.3)
    cd = O.complex_view(O.ic_delta(c, 5001, pk, sphere, 1, 0), n)
    s = signed(n)
    kx, ky, kz = s[:, None, None], s[None, :, None], np.arange(n // 2 + 1)[None, None, :]
    n2 = kx ** 2 + ky ** 2 + kz ** 2
    nyq = (np.abs(kx) == n // 2) | (np.abs(ky) == n // 2) | (kz == n // 2)
    if sphere:
        keep = n2 <= (ns // 2) ** 2
    else:
        keep = (np.abs(kx) <= ns // 2) & (np.abs(ky) <= ns // 2) & (kz <= ns // 2)
    keep = keep & ~nyq & (n2 > 0)
    amp2 = np.abs(cd) ** 2
    np.testing.assert_allclose(amp2[keep], pk[n2[keep]] / box ** 3, rtol=1e-12)
    assert np.all(amp2[~keep] == 0.0)
    # Hermitian kz = 0 plane: delta(-kx, -ky, 0) = conj delta(kx, ky, 0)
    p0 = cd[:, :, 0]
    mirror = p0[(-np.arange(n)) % n][:, (-np.arange(n)) % n]
    np.testing.assert_allclose(p0, np.conj(mirror), atol=0)
    # inverted initial conditions: exactly -delta
    inv = O.complex_view(O.ic_delta(c, 5001, pk, sphere, 1, 1), n)
    assert np.array_equal(inv, -cd)


def test_ic_delta_draws_consumed_by_skipped_modes():
    """The cut to |n_a| <= Nsample/2 (src/2LPT.c:380-386) only removes modes:
    the modes both cuts keep carry the same numbers for Nsample = 16 and
    Nsample = 8 (the IC of a coarser particle grid is the finer one's, cut)."""
    n, box = 16, 48.0
    pk = ic_pk(n, box)
    big = O.complex_view(O.ic_delta(O.cfg(n, 16, box, 0.3), 5001, pk, 0, 0, 0), n)
    small = O.complex_view(O.ic_delta(O.cfg(n, 8, box, 0.3), 5001, pk, 0, 0, 0), n)
    nz = small != 0
    assert nz.sum() > 100
    assert np.array_equal(small[nz], big[nz])


@pytest.mark.parametrize("amp_fixed", [0, 1])
def test_ic_delta_phase_is_column_stream_draw(amp_fixed):
    """The phase of mode (i, j, k), k > 0, is 2 pi times draw number k (amplitude
    fixed) or 2k (phase and amplitude drawn per mode) of the stream seeded with
    the seed table entry of column (i, j) (src/2LPT.c:337-357); without fixed
    amplitudes the modulus is sqrt(-ln u P / Box^3) with u the next draw.  The
    k = 0 mode of column (0, 0) is skipped but still draws (src/2LPT.c:347-357),
    so that column's modes k > 0 keep their place in the stream."""
    n, box = 12, 36.0
    pk = ic_pk(n, box)
    cd = O.complex_view(O.ic_delta(O.cfg(n, n, box, 0.3), 5001, pk, 0, amp_fixed, 0), n)
    st = O.seedtable(n, 5001).reshape(n, n)
    s = signed(n)
    for (i, j, k) in [(1, 2, 3), (5, 0, 1), (0, 7, 4), (11, 11, 2), (3, 9, 5), (0, 0, 1), (0, 0, 4)]:
        u = O.ranlxd_stream(int(st[i, j]), 2 * (k + 1))
        if amp_fixed:
            ph, mod2 = 2 * np.pi * u[k], pk[s[i] ** 2 + s[j] ** 2 + k * k] / box ** 3
        else:
            ph = 2 * np.pi * u[2 * k]
            mod2 = -np.log(u[2 * k + 1]) * pk[s[i] ** 2 + s[j] ** 2 + k * k] / box ** 3
        v = cd[i, j, k]
        assert abs(np.angle(v) - np.angle(np.exp(1j * ph))) < 1e-9, (i, j, k)
        assert abs(abs(v) ** 2 - mod2) < 1e-9 * mod2, (i, j, k)


def test_ic_delta_amplitudes_are_exponential():
    """|delta|^2 Box^3 / P = -ln u is exponentially distributed with mean 1
    (Rayleigh-distributed amplitudes of a Gaussian field)."""
    n, box = 32, 96.0
    pk = ic_pk(n, box)
    cd = O.complex_view(O.ic_delta(O.cfg(n, n, box, 0.3), 5001, pk, 0, 0, 0), n)
    s = signed(n)
    n2 = s[:, None, None] ** 2 + s[None, :, None] ** 2 + np.arange(n // 2 + 1)[None, None, :] ** 2
    nz = np.abs(cd) > 0
    nz[:, :, 0] = False  # the kz = 0 plane holds each draw twice (Hermitian copies)
    x = (np.abs(cd[nz]) ** 2 * box ** 3 / pk[n2[nz]])
    assert x.size > 5000
    assert abs(x.mean() - 1.0) < 4.0 / np.sqrt(x.size)
    # Kolmogorov-Smirnov distance to 1 - exp(-x)
    xs = np.sort(x)
    ks = np.max(np.abs(np.arange(1, xs.size + 1) / xs.size - (1 - np.exp(-xs))))
    assert ks < 1.6 / np.sqrt(xs.size)


# ---------------------------------------------------------------- LPT fields (scale-dependent build)

def test_lpt_field_with_unit_growth_is_za():
    n, ns, box = 16, 8, 48.0
    c = O.cfg(n, ns, box, 0.3)
    cd = O.ic_delta(c, 5001, ic_pk(n, box))
    za, lp = O.ic_displacements(c, cd, 0)
    g1 = np.ones(3 * (n // 2) ** 2 + 1)
    np.testing.assert_allclose(O.lpt_field(c, cd, g1), za, rtol=0, atol=1e-6 * np.abs(za).max())
    (cd2,) = O.ic_displacements(c, cd, 1)
    g2 = np.full_like(g1, -3.0 / 7.0 / n ** 3)
    np.testing.assert_allclose(O.lpt_field(c, cd2, g2), lp, rtol=0, atol=1e-6 * np.abs(lp).max())
    assert np.abs(lp).max() > 1e-3 * np.abs(za).max()


def test_lpt_field_plane_waves_with_scale_dependent_growth():
    """Two waves of different |k| with growth g(|n|^2): each displacement is
    its own ZA field scaled by g of its wavenumber, read at the lattice
    (N = 2 Nsample, so the lattice points are mesh points)."""
    n, ns, box = 16, 8, 64.0
    A, B = 0.05, 0.03
    m1, m2 = (0, 0, 2), (2, 2, 0)
    cd = wave_spectrum(n, [(A, m1), (B, m2)])
    n2 = np.arange(3 * (n // 2) ** 2 + 1)
    g = 1.0 + 0.1 * n2  # g(4) = 1.4, g(8) = 1.8
    out = O.lpt_field(O.cfg(n, ns, box, 0.3), cd, g)
    qi = np.indices((ns, ns, ns)).reshape(3, -1) * (n // ns)
    kf = 2 * np.pi / box
    expect = np.zeros((3, ns ** 3))
    for amp, m in ((A, m1), (B, m2)):
        kv = kf * np.array(m, float)
        th = 2 * np.pi * (np.array(m)[:, None] * qi).sum(0) / n
        gm = g[int(np.dot(m, m))]
        for a in range(3):
            expect[a] += -amp * gm * kv[a] / kv.dot(kv) * np.sin(th)
    expect -= expect.mean(axis=1, keepdims=True)
    np.testing.assert_allclose(out, expect, atol=1e-7)
