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
  Kick / Drift           free streaming with zero force, periodic wrap
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
