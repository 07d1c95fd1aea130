"""C host (libmgpicola.so) on the CPU: paramfile interface, growth factors and
COLA time integrals, model hooks, power spectrum normalisation and the output
formats.  No GPU is used: these functions run before and around the device path.

Fixtures: the reference's example paramfiles, its output-redshift list and its
tabulated input power spectrum (tests/golden/).  Known answers are closed
forms (Einstein-de Sitter growth and time integrals, the hook formulas, sigma8
of the normalised spectrum) or independent numerical solutions (scipy).
"""
import ctypes as C
import os
import re

import numpy as np
import pytest
from scipy import integrate

import mgpicola as M

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
PARAMFILES = {
    "example_fofr_withscreening.txt": "FOFR",
    "example_fofr_noscreening.txt": "FOFR",
    "example_fofr_withscreening_lcdmgrowthfac.txt": "FOFR",
    "example_lcdm_fofrversion.txt": "FOFR",
    "example_dgp_withscreening.txt": "DGP",
    "example_dgp_noscreening.txt": "DGP",
    "example_lcdm_dgpversion.txt": "DGP",
    "example_mbeta.txt": "MBETA",
}


@pytest.fixture
def in_golden(monkeypatch):
    monkeypatch.chdir(GOLDEN)


def tag_values(path):
    """tag -> value text of a paramfile (first token after the tag, % starts a comment)."""
    out = {}
    for line in open(path):
        line = line.split("%")[0].strip()
        if not line:
            continue
        parts = line.split()
        if len(parts) >= 2:
            out[parts[0]] = parts[1]
    return out


@pytest.mark.parametrize("fname", sorted(PARAMFILES))
def test_reference_paramfiles_parse(in_golden, fname):
    model = PARAMFILES[fname]
    hp = M.read_params(os.path.join("paramfiles", fname), model=model, compute_pofk=False)
    for tag, val in tag_values(os.path.join("paramfiles", fname)).items():
        got = getattr(hp, tag)
        if isinstance(got, bytes):
            assert got.decode() == val, tag
        else:
            assert float(got) == pytest.approx(float(val), rel=1e-15), tag
    # the output list: (z, nsteps) pairs sorted by decreasing redshift
    rows = [l for l in open("files/output_redshifts.dat") if l.strip() and not l.startswith("%")]
    ref = sorted(((float(r.split(",")[0]), int(r.split(",")[1])) for r in rows), key=lambda t: -t[0])
    assert hp.Noutputs == len(ref)
    for i, (z, n) in enumerate(ref):
        assert hp.out_z[i] == pytest.approx(z) and hp.out_nsteps[i] == n


def test_paramfile_is_model_specific(in_golden):
    with pytest.raises(ValueError, match="fofr0"):
        M.read_params("paramfiles/example_fofr_withscreening.txt", model="DGP", compute_pofk=False)


def test_bransdicke_example_tag_names(in_golden, tmp_path):
    """The reference registers the JBD tags as 'Omegar2' and 'Omegav2'
    (src/user_defined_functions.h:393-398) while its example file spells them
    'Omegarh2'/'Omegavh2': the strict reader rejects the example as shipped,
    and accepts it with the registered names.  The example leaves the JBD
    values blank, so the token after the tag is the comment's '%', which the
    reference converts with atof to 0 (src/read_param.c:376-379)."""
    src = "paramfiles/example_bransdicke.txt"
    with pytest.raises(ValueError, match="Omegarh2"):
        M.read_params(src, model="BRANSDICKE", compute_pofk=False)
    text = open(src).read().replace("Omegarh2", "Omegar2 ").replace("Omegavh2", "Omegav2 ")
    fixed = tmp_path / "bd.txt"
    fixed.write_text(text.replace("files/output_redshifts.dat", os.path.join(GOLDEN, "files/output_redshifts.dat"))
                     .replace("files/input_power_spectrum.dat", os.path.join(GOLDEN, "files/input_power_spectrum.dat")))
    hp = M.read_params(str(fixed), model="BRANSDICKE", compute_pofk=False)
    assert (hp.wBD, hp.Omegah2, hp.Omegarh2, hp.Omegavh2) == (0.0, 0.0, 0.0, 0.0)
    assert hp.Nmesh == int(tag_values(src)["Nmesh"])


def _write(tmp_path, text):
    p = tmp_path / "p.txt"
    p.write_text(text)
    return str(p)


def test_paramfile_strictness(in_golden, tmp_path):
    base = open("paramfiles/example_fofr_withscreening.txt").read()
    # a missing tag
    missing = re.sub(r"^Seed .*$", "", base, flags=re.M)
    with pytest.raises(ValueError, match="I miss a value for tag 'Seed'"):
        M.read_params(_write(tmp_path, missing), model="FOFR", compute_pofk=False)
    # a duplicated tag
    with pytest.raises(ValueError, match="Seed"):
        M.read_params(_write(tmp_path, base + "\nSeed 7\n"), model="FOFR", compute_pofk=False)
    # an unknown tag
    with pytest.raises(ValueError, match="NotATag"):
        M.read_params(_write(tmp_path, base + "\nNotATag 1\n"), model="FOFR", compute_pofk=False)
    # the P(k) tags are required exactly when P(k) is computed
    with pytest.raises(ValueError, match="pofk_nbins"):
        M.read_params(_write(tmp_path, base), model="FOFR", compute_pofk=True)


def test_output_list_errors(in_golden, tmp_path):
    base = open("paramfiles/example_fofr_withscreening.txt").read()
    bad = tmp_path / "out.dat"
    bad.write_text("0.5\n")
    pf = _write(tmp_path, base.replace("files/output_redshifts.dat", str(bad)))
    with pytest.raises(ValueError, match="incorrect format"):
        M.read_params(pf, model="FOFR", compute_pofk=False)
    bad.write_text("1.0, 0\n")
    with pytest.raises(ValueError, match="nsteps of <= 0"):
        M.read_params(pf, model="FOFR", compute_pofk=False)


# ---------------------------------------------------------------- cosmology

def cosmo_for(tmp_path, model="FOFR", **over):
    vals = {}
    with open(os.path.join(GOLDEN, "paramfiles", "example_fofr_withscreening.txt" if model.startswith("FOFR") else
                           {"DGP": "example_dgp_withscreening.txt", "MBETA": "example_mbeta.txt"}[model])) as f:
        for line in f:
            s = line.split("%")[0].split()
            if len(s) >= 2:
                vals[s[0]] = s[1]
    vals["OutputRedshiftFile"] = os.path.join(GOLDEN, "files/output_redshifts.dat")
    vals["FileWithInputSpectrum"] = os.path.join(GOLDEN, "files/input_power_spectrum.dat")
    vals["OutputDir"] = str(tmp_path)
    vals.update({k: str(v) for k, v in over.items()})
    pf = tmp_path / "p.txt"
    pf.write_text("".join(f"{k} {v}\n" for k, v in vals.items()))
    hp = M.read_params(str(pf), model=model, compute_pofk=False)
    hp.growth_on_gpu = 0  # the host's own ODE solver: these tests run without a GPU
    return hp, M.Cosmo(hp)


def test_eds_growth_and_time_integrals(tmp_path):
    """Omega = 1: D1 = a, D2 = a^2 (normalised at a = 1), dD/dy = Q dD/da with Q = a^3 H = a^1.5,
    Sq = (ai^-3 - af^-3)/3 * aRef^2.5, Sphi = -(af^-2.5 - ai^-2.5) aRef^3 / 2.5."""
    hp, c = cosmo_for(tmp_path, model="FOFR", modified_gravity_active=0, Omega=1.0, OmegaBaryon=0.04)
    for a in (0.05, 0.1, 0.3, 0.7, 1.0):
        assert c.growth(0, a) == pytest.approx(a, rel=2e-6)
        assert c.growth(3, a) == pytest.approx(a * a, rel=2e-5)
        assert c.growth(1, a) == pytest.approx(a ** 1.5, rel=2e-5)
        assert c.growth(2, a) == pytest.approx(1.5 * a * a, rel=2e-5)
    for ai, af, ar in ((0.1, 0.13, 0.115), (0.5, 0.6, 0.6), (0.05, 1.0, 0.5)):
        assert c.Sq(ai, af, ar) == pytest.approx((ai ** -3 - af ** -3) / 3 * ar ** 2.5, rel=1e-9)
        assert c.Sphi(ai, af, ar) == pytest.approx(-(af ** -2.5 - ai ** -2.5) * ar ** 3 / 2.5, rel=1e-12)
    c.close()


def lcdm_growth_exact(om, a):
    """Growing mode D+ ~ H(a) int_0^a da' / (a' H(a'))^3, normalised at a = 1."""
    H = lambda x: np.sqrt(om / x ** 3 + 1 - om)
    f = lambda x: 1.0 / (x * H(x)) ** 3
    d = lambda x: H(x) * integrate.quad(f, 0, x, epsabs=0, epsrel=1e-12)[0]
    return d(a) / d(1.0)


def lcdm_d2_ode(om, a_eval, zini):
    """Second-order growth from an independent integration: D2'' + (2 + a H'/H) D2' = 1.5 Om/(a^3 H^2) (D2 - D1^2) in x = ln a."""
    def rhs(x, y):
        a = np.exp(x)
        H = np.sqrt(om / a ** 3 + 1 - om)
        dH = -1.5 * om / a ** 4 / H
        al, be = 2 + a * dH / H, 1.5 * om / (a ** 3 * H * H)
        return [y[1], -al * y[1] + be * y[0], y[3], -al * y[3] + be * (y[2] - y[0] ** 2)]
    x0 = np.log(1 / (1 + zini))
    xs = np.log(np.append(a_eval, 1.0))
    s = integrate.solve_ivp(rhs, (x0, 0.0), [1, 1, -3 / 7, -6 / 7], t_eval=xs, method="DOP853", rtol=1e-11, atol=1e-13)
    return s.y[0][:-1] / s.y[0][-1], s.y[2][:-1] / s.y[2][-1]


def test_lcdm_growth_against_exact_solution(tmp_path):
    hp, c = cosmo_for(tmp_path, model="FOFR", modified_gravity_active=0)
    om = hp.Omega
    for a in (0.05, 0.2, 0.5, 0.8):
        assert c.growth(0, a) == pytest.approx(lcdm_growth_exact(om, a), rel=1e-5)
        # d^2D/dy^2 = 1.5 Om a D
        assert c.growth(2, a) == pytest.approx(1.5 * om * a * c.growth(0, a), rel=1e-6)
        # dD/dy = Q dD/da, Q = a^3 H
        h = 1e-4 * a
        dDda = (lcdm_growth_exact(om, a + h) - lcdm_growth_exact(om, a - h)) / (2 * h)
        Q = a ** 3 * np.sqrt(om / a ** 3 + 1 - om)
        assert c.growth(1, a) == pytest.approx(Q * dDda, rel=1e-5)
    a_eval = np.array([0.05, 0.2, 0.5, 0.8])
    d1, d2 = lcdm_d2_ode(om, a_eval, 200.0)  # the tables start at z = max(200, Init_Redshift)
    for i, a in enumerate(a_eval):
        assert c.growth(3, a) == pytest.approx(d2[i], rel=1e-4)
    c.close()


def test_fofr_hooks_closed_forms(tmp_path):
    hp, c = cosmo_for(tmp_path, model="FOFR")
    om, f0, n = hp.Omega, hp.fofr0, hp.nfofr
    for a in (0.1, 0.5, 1.0):
        fac, fac0 = om / a ** 3 + 4 * (1 - om), om + 4 * (1 - om)
        assert c.mass2(a) == pytest.approx(fac0 * (fac / fac0) ** (n + 2) / ((1 + n) * f0), rel=1e-13)
        assert c.coupling(a) == pytest.approx(1.0 / 3.0, rel=1e-15)
        assert c.phi_crit(a) == pytest.approx(1.5 * f0 * (fac0 / fac) ** (n + 1), rel=1e-13)
        k = 0.3
        k2 = (k * 2997.92458) ** 2
        assert c.GeffoverG(a, k) == pytest.approx(1 + k2 / (3 * (k2 + a * a * c.mass2(a))), rel=1e-13)
        m = c.mg(a)
        assert (m.coupling, m.mass2, m.phi_crit) == pytest.approx((c.coupling(a), c.mass2(a), c.phi_crit(a)))
    c.close()


def test_dgp_hooks_closed_forms(tmp_path):
    hp, c = cosmo_for(tmp_path, model="DGP")
    om, rc = hp.Omega, hp.rcH0_DGP
    for a in (0.2, 0.6, 1.0):
        H = np.sqrt(om / a ** 3 + 1 - om)
        dH = -1.5 * om / a ** 4 / H
        beta = 1 + 2 * rc * (H + a * dH / 3)
        assert c.beta_dgp(a) == pytest.approx(beta, rel=1e-13)
        assert c.coupling(a) == pytest.approx(1 / (3 * beta), rel=1e-13)
    c.close()


def test_fofr_scale_dependent_growth_limits(tmp_path):
    """k -> 0: D(k, a) -> D_LCDM(a); at finite k the table equals an independent
    ODE solution with mu(k, a) = 1 + (1/3) k^2 / (k^2 + a^2 m^2(a))."""
    hp, c = cosmo_for(tmp_path, model="FOFR")
    om = hp.Omega
    a_eval = np.array([0.1, 0.4, 1.0])
    for a in a_eval[:2]:  # the smallest tabulated k is pi / Box
        assert c.growth_sd(0, 0.02, a) == pytest.approx(c.growth_lcdm(0, a), rel=5e-3)

    def rhs(x, y, k):
        a = np.exp(x)
        H = np.sqrt(om / a ** 3 + 1 - om)
        dH = -1.5 * om / a ** 4 / H
        k2 = (k * 2997.92458) ** 2
        mu = 1 + k2 / (3 * (k2 + a * a * c.mass2(a)))
        return [y[1], -(2 + a * dH / H) * y[1] + 1.5 * om * mu / (a ** 3 * H * H) * y[0]]

    for k in (0.02, 0.05, 0.5, 2.0):
        a0 = 1 / (1 + 200.0)
        s = integrate.solve_ivp(rhs, (np.log(a0), 0.0), [1, 1], t_eval=np.log(a_eval), args=(k,), method="DOP853",
                                rtol=1e-11, atol=1e-13)
        d = s.y[0] / s.y[0][-1]
        for i, a in enumerate(a_eval):
            assert c.growth_sd(0, k, a) == pytest.approx(d[i], rel=2e-4), (k, a)
    c.close()


def test_sigma8_normalisation(tmp_path):
    """sigma8 of the normalised linear P(k) at z = 0 equals the paramfile's Sigma8 (LCDM version)."""
    hp, c = cosmo_for(tmp_path, model="FOFR", modified_gravity_active=0)

    def w(x):
        return 3 * (np.sin(x) - x * np.cos(x)) / x ** 3

    f = lambda lk: (np.exp(lk) ** 3 * c.power_spec(np.exp(lk)) * w(8 * np.exp(lk)) ** 2 / (2 * np.pi ** 2))
    s2 = integrate.quad(f, np.log(1e-5), np.log(500 / 8.0), limit=400, epsrel=1e-9)[0]
    assert np.sqrt(s2) == pytest.approx(hp.Sigma8, rel=2e-3)
    c.close()


def growth_ode(om, mu, a_eval, zini=200.0):
    """Raw growing mode from D = 1, dD/dln a = 1 at z_ini (the tables' start), independent of the host:
    D'' + (2 + a H'/H) D' = 1.5 Om mu(a) / (a^3 H^2) D in x = ln a."""
    def rhs(x, y):
        a = np.exp(x)
        H = np.sqrt(om / a ** 3 + 1 - om)
        dH = -1.5 * om / a ** 4 / H
        return [y[1], -(2 + a * dH / H) * y[1] + 1.5 * om * mu(a) / (a ** 3 * H * H) * y[0]]
    x0 = np.log(1 / (1 + zini))
    s = integrate.solve_ivp(rhs, (x0, 0.0), [1.0, 1.0], t_eval=np.log(a_eval), method="DOP853", rtol=1e-11,
                            atol=1e-13)
    return s.y[0]


def ic_table(tmp_path, model, **over):
    hp, c = cosmo_for(tmp_path, model=model, Nmesh=32, Nsample=16, Box=128.0, **over)
    t = np.zeros(M.n2_table_len(32))
    M.hlib().mgp_ic_pk_table(c.h, t.ctypes.data)
    return hp, c, t


def test_mg_pofk_ratio_scale_independent(tmp_path):
    """nDGP, input P(k) for LCDM (src/2LPT.c:329-410, src/cosmo.c:688-756): with Sigma8
    given for LCDM the IC power is P_LCDM (D_DGP(1)/D_LCDM(1))^2, the ratio of raw growing
    modes from the same start (independent ODE); with Sigma8 given for the MG run the
    ratio and the sigma8 enhancement cancel and the table is the unrescaled one."""
    hp, c0, t0 = ic_table(tmp_path, "DGP", modified_gravity_active=1, input_pofk_is_for_lcdm=0)
    rc = hp.rcH0_DGP
    om = hp.Omega

    def mu(a):
        H = np.sqrt(om / a ** 3 + 1 - om)
        dH = -1.5 * om / a ** 4 / H
        return 1 + 1 / (3 * (1 + 2 * rc * (H + a * dH / 3)))

    ratio = (growth_ode(om, mu, [1.0])[0] / growth_ode(om, lambda a: 1.0, [1.0])[0]) ** 2
    _, c1, t1 = ic_table(tmp_path, "DGP", modified_gravity_active=1, input_pofk_is_for_lcdm=1,
                         input_sigma8_is_for_lcdm=1)
    assert ratio > 1.01
    np.testing.assert_allclose(t1[1:] / t0[1:], ratio, rtol=2e-6)
    assert c1.mg_pofk_ratio(0.1, 1.0) == pytest.approx(ratio, rel=2e-6)
    assert c1.mg_sigma8_enhancement(1.0) == pytest.approx(np.sqrt(ratio), rel=2e-6)
    _, c2, t2 = ic_table(tmp_path, "DGP", modified_gravity_active=1, input_pofk_is_for_lcdm=1,
                         input_sigma8_is_for_lcdm=0)
    np.testing.assert_allclose(t2, t0, rtol=1e-12)
    for c in (c0, c1, c2):
        c.close()


def test_mg_pofk_ratio_scale_dependent(tmp_path):
    """f(R) (scale-dependent): the ratio is (D(k, 1)/D_LCDM(1))^2 with D(k) the independent
    ODE solution for mu(k, a) = 1 + k^2 / (3 (k^2 + a^2 m^2)), 1 on large scales; with Sigma8
    given for the MG run the rescaled spectrum keeps sigma8 = Sigma8 (the enhancement is the
    sigma8 ratio of the reference's 1000-point integral)."""
    hp, c, t = ic_table(tmp_path, "FOFR", modified_gravity_active=1, input_pofk_is_for_lcdm=1,
                        input_sigma8_is_for_lcdm=1)
    om = hp.Omega
    dl = growth_ode(om, lambda a: 1.0, [1.0])[0]
    for k in (0.05, 0.3, 1.0):
        k2 = (k * 2997.92458) ** 2
        dk = growth_ode(om, lambda a: 1 + k2 / (3 * (k2 + a * a * c.mass2(a))), [1.0])[0]
        assert c.mg_pofk_ratio(k, 1.0) == pytest.approx((dk / dl) ** 2, rel=5e-4), k
    # towards 1 on large scales (the lowest tabulated k is half the fundamental mode)
    assert 1.0 < c.mg_pofk_ratio(0.025, 1.0) < 1.01 < c.mg_pofk_ratio(0.3, 1.0) < c.mg_pofk_ratio(1.0, 1.0)
    kf = 2 * np.pi / hp.Box
    for n2 in (1, 50, 700):
        k = kf * np.sqrt(n2)
        assert t[n2] == pytest.approx(c.power_spec(k) * c.mg_pofk_ratio(k, 1.0), rel=1e-12)
    s8 = c.mg_sigma8_enhancement(1.0)
    kk = np.exp(np.linspace(np.log(1e-3), np.log(100.0), 1000))
    w = 3 * (np.sin(8 * kk) - 8 * kk * np.cos(8 * kk)) / (8 * kk) ** 3
    pw = np.array([c.power_spec(k) for k in kk])
    rt = np.array([c.mg_pofk_ratio(k, 1.0) for k in kk])
    s_mg = np.sqrt(np.sum(pw * rt * w * w * kk ** 3) / np.sum(pw * w * w * kk ** 3))
    assert s8 == pytest.approx(s_mg, rel=1e-4)
    assert s8 > 1.01
    c.close()
    _, c2, t2 = ic_table(tmp_path, "FOFR", modified_gravity_active=1, input_pofk_is_for_lcdm=1,
                         input_sigma8_is_for_lcdm=0)
    np.testing.assert_allclose(t2[1:], t[1:] / s8 ** 2, rtol=1e-12)
    c2.close()


# ---------------------------------------------------------------- output formats

def test_gadget_snapshot_layout(tmp_path):
    hp, c = cosmo_for(tmp_path, model="FOFR", Nsample=4, Nmesh=8)
    c.close()
    L = M.hlib()
    L.mgp_write_gadget.argtypes = [C.POINTER(M.HostParams), C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_size_t]
    npart = 64
    rng = np.random.default_rng(0)
    pos = rng.uniform(0, hp.Box, size=(3, npart)).astype(np.float32)
    vel = rng.normal(size=(3, npart)).astype(np.float32)
    A = 0.5
    assert L.mgp_write_gadget(C.byref(hp), A, 1.0, pos.ctypes.data, vel.ctypes.data, npart) == 0
    fn = tmp_path / f"{hp.FileBase.decode()}_z1p000.0"
    raw = open(fn, "rb").read()
    assert len(raw) == 4 + 256 + 4 + 2 * (4 + 12 * npart + 4)
    b0, = np.frombuffer(raw[:4], np.int32)
    assert b0 == 256
    h = raw[4:260]
    npart6 = np.frombuffer(h[0:24], np.uint32)
    mass = np.frombuffer(h[24:72], np.float64)
    t, z = np.frombuffer(h[72:88], np.float64)
    ntot = np.frombuffer(h[96:120], np.uint32)
    nfiles, = np.frombuffer(h[124:128], np.int32)
    box, om0, ol, hh = np.frombuffer(h[128:160], np.float64)
    assert npart6[1] == npart and ntot[1] == hp.Nsample ** 3 and nfiles == 1
    assert t == A and z == pytest.approx(1 / A - 1)
    assert box == pytest.approx(hp.Box * hp.UnitLength_in_cm / 3.085678e24)
    assert (om0, ol, hh) == pytest.approx((hp.Omega, 1 - hp.Omega, hp.HubbleParam))
    # particle mass: rho_crit Omega Box^3 / N in code units
    ut = hp.UnitLength_in_cm / hp.UnitVelocity_in_cm_per_s
    G = 6.672e-8 / hp.UnitLength_in_cm ** 3 * hp.UnitMass_in_g * ut ** 2
    H = 3.2407789e-18 * ut
    assert mass[1] == pytest.approx(3 * hp.Omega * H * H * hp.Box ** 3 / (8 * np.pi * G * hp.Nsample ** 3), rel=1e-12)
    off = 264
    assert np.frombuffer(raw[off - 4:off], np.int32)[0] == 256
    assert np.frombuffer(raw[off:off + 4], np.int32)[0] == 12 * npart
    p = np.frombuffer(raw[off + 4:off + 4 + 12 * npart], np.float32).reshape(npart, 3)
    np.testing.assert_array_equal(p, (pos.T * np.float32(hp.UnitLength_in_cm / 3.085678e24)).astype(np.float32))
    off2 = off + 4 + 12 * npart + 4
    v = np.frombuffer(raw[off2 + 4:off2 + 4 + 12 * npart], np.float32).reshape(npart, 3)
    np.testing.assert_array_equal(v, vel.T)


def test_pofk_file_format(tmp_path):
    hp, c = cosmo_for(tmp_path, model="FOFR", Nsample=16, Nmesh=16)
    c.close()
    L = M.hlib()
    dp = C.POINTER(C.c_double)
    L.mgp_write_pofk.argtypes = [C.POINTER(M.HostParams), C.c_double, C.c_char_p, C.c_void_p, C.c_void_p, C.c_void_p,
                                 C.c_int, C.c_int, C.c_double, C.c_double]
    nb = 8
    pk = np.linspace(1, 8, nb) * 100
    km = np.linspace(0.01, 0.2, nb)
    nm = np.array([0, 4, 6, 0, 10, 12, 20, 30], dtype=float)
    assert L.mgp_write_pofk(C.byref(hp), 0.5, b"cb", pk.ctypes.data, km.ctypes.data, nm.ctypes.data, nb, 0, 0.0,
                            16.0) == 0
    fn = tmp_path / f"pofk_{hp.FileBase.decode()}_z1.000_cb.txt"
    lines = open(fn).read().splitlines()
    assert lines[0].startswith("#")
    rows = np.array([[float(x) for x in l.split()] for l in lines[1:]])
    keep = [i for i in range(1, nb) if nm[i] > 0]  # bin 0 is never written
    assert rows.shape == (len(keep), 4)
    kb = np.array([2 * np.pi / hp.Box * (i / nb * 16.0) for i in keep])
    np.testing.assert_allclose(rows[:, 0], kb, atol=6e-6)
    np.testing.assert_allclose(rows[:, 1], pk[keep], atol=6e-6)
    np.testing.assert_allclose(rows[:, 2], km[keep], atol=6e-6)
    np.testing.assert_allclose(rows[:, 3], pk[keep] * kb ** 3 / (2 * np.pi ** 2), rtol=1e-4, atol=6e-6)


# ---------------------------------------------------------------- massive neutrinos (FOFRNU)
CAMB = os.path.join(GOLDEN, "camb_nu0.2")
CAMB_Z = ["0.000", "0.494", "1.059", "2.075", "4.393", "10.103", "20.098", "50.000"]


def camb_info(tmp_path):
    """An info file in the reference's format (directory and count, then file and z)
    listing the fixture subset of the reference's nu0.2 CAMB tables."""
    p = tmp_path / "transfer_info.txt"
    p.write_text(f"{CAMB} {len(CAMB_Z)}\n" + "".join(f"lcdm_nu0.2_transfer_z{z}.dat {z}\n" for z in CAMB_Z))
    return str(p)


def nu_cosmo(tmp_path, **over):
    vals = tag_values(os.path.join(GOLDEN, "paramfiles", "fofr_nu0.2.inp"))
    vals["nu_FilenameTransferInfofile"] = camb_info(tmp_path)
    vals["OutputRedshiftFile"] = os.path.join(GOLDEN, "files/output_redshifts.dat")
    vals["OutputDir"] = str(tmp_path)
    vals.update({k: str(v) for k, v in over.items()})
    pf = tmp_path / "p.txt"
    pf.write_text("".join(f"{k} {v}\n" for k, v in vals.items()))
    hp = M.read_params(str(pf), model="FOFRNU", compute_pofk=True)
    hp.growth_on_gpu = 0  # the host's own ODE solver: these tests run without a GPU
    return hp, M.Cosmo(hp)


def camb_rows(z):
    return np.loadtxt(os.path.join(CAMB, f"lcdm_nu0.2_transfer_z{z}.dat"), skiprows=1)


def test_fofrnu_paramfile_and_omeganu(in_golden, tmp_path):
    """The reference's f(R) + massive-nu paramfile parses as FOFRNU (its tag set,
    with the output list moved to a file that exists here); OmegaNu =
    SumMassNuEV / (93.14 h^2), OmegaCDM = Omega - OmegaBaryon - OmegaNu."""
    src = open(os.path.join("paramfiles", "fofr_nu0.2.inp")).read().splitlines()
    out = os.path.join(GOLDEN, "files/output_redshifts.dat")
    pf = tmp_path / "fofr_nu0.2.inp"
    pf.write_text("\n".join(f"OutputRedshiftFile {out}" if l.split("%")[0].split()[:1] == ["OutputRedshiftFile"]
                             else l for l in src) + "\n")
    hp = M.read_params(str(pf), model="FOFRNU", compute_pofk=True)
    assert hp.nu_include_massive_neutrinos == 1 and hp.nu_SumMassNuEV == 0.2
    with pytest.raises(ValueError, match="nu_SumMassNuEV"):
        M.read_params(os.path.join("paramfiles", "example_fofr_withscreening.txt"), model="FOFRNU",
                      compute_pofk=False)
    hp, c = nu_cosmo(tmp_path, Nmesh=16, Nsample=16)
    h = hp.HubbleParam
    assert hp.OmegaNu == pytest.approx(0.2 / (93.14 * h * h), rel=1e-14)
    assert hp.OmegaCDM == pytest.approx(hp.Omega - hp.OmegaBaryon - hp.OmegaNu, rel=1e-14)
    c.close()


def test_camb_transfer_functions_at_the_table_nodes(tmp_path):
    """The (ln k, z) splines pass through the CAMB values at every node; T_cb
    weighs CDM and baryons by their densities."""
    hp, c = nu_cosmo(tmp_path, Nmesh=16, Nsample=16)
    for z in ("0.000", "4.393"):
        r = camb_rows(z)
        a = 1.0 / (1.0 + float(z))
        for row in r[::17]:
            k = row[0]
            assert c.T_nu(k, a) == pytest.approx(row[5], rel=1e-10)
            assert c.T_cdm(k, a) == pytest.approx(row[1], rel=1e-10)
            assert c.T_b(k, a) == pytest.approx(row[2], rel=1e-10)
            assert c.T_tot(k, a) == pytest.approx(row[6], rel=1e-10)
            cb = (hp.OmegaBaryon * row[2] + hp.OmegaCDM * row[1]) / (hp.OmegaBaryon + hp.OmegaCDM)
            assert c.T_cb(k, a) == pytest.approx(cb, rel=1e-10)
    c.close()


def test_munu_limits(tmp_path):
    """mu_nu (1LPT) -> 1 on scales where neutrinos cluster like CDM and ->
    (Omega - OmegaNu)/Omega where they free-stream; the 2LPT factor is the latter."""
    hp, c = nu_cosmo(tmp_path, Nmesh=16, Nsample=16)
    f = (hp.Omega - hp.OmegaNu) / hp.Omega
    for a in (0.05, 0.5, 1.0):
        assert c.munu_2lpt(a, 0.3) == f
        assert c.munu_1lpt(a, 1e-4) == pytest.approx(1.0, abs=2e-3)
        assert c.munu_1lpt(a, 5.0) == pytest.approx(f, abs=2e-3)
        assert f < c.munu_1lpt(a, 0.1) < 1.0
    c.close()


def test_camb_total_power_matches_camb_matter_power(tmp_path):
    """CAMB's own z = 0 matter P(k) over P_total(k) = 2 pi^2 T_tot^2 (kh) (kh/0.05)^(ns-1) h^3
    is the primordial amplitude (~2.2e-9) times at most a slow power of k (the CAMB
    run's tilt need not be the paramfile's) over three decades: a wrong column or a
    wrong k unit would leave structure in the ratio."""
    hp, c = nu_cosmo(tmp_path, Nmesh=16, Nsample=16)
    mp = np.loadtxt(os.path.join(CAMB, "lcdm_nu0.2_matterpower_z0.000.dat"), comments="#")
    sel = (mp[:, 0] > 1e-3) & (mp[:, 0] < 1.0)
    ratio = mp[sel, 1] / np.array([c.pk_total(k) for k in mp[sel, 0]])
    lk, lr = np.log(mp[sel, 0]), np.log(ratio)
    slope, icpt = np.polyfit(lk, lr, 1)
    assert abs(slope) < 0.01 and 1.5e-9 < np.exp(icpt + slope * np.log(0.05)) < 3e-9, (slope, icpt)
    assert np.std(lr - (slope * lk + icpt)) < 5e-3
    # and P_cb = P_total (T_cb / T_tot)^2
    for k in (0.01, 0.1, 1.0):
        assert c.pk_cb(k) == pytest.approx(c.pk_total(k) * (c.T_cb(k, 1.0) / c.T_tot(k, 1.0)) ** 2, rel=1e-13)
    c.close()


def test_fofrnu_growth_matches_ode(tmp_path):
    """The scale-dependent D(k, a) with neutrinos equals an independent ODE solution
    with mu = mu_fofr(k, a) mu_nu(k, a) (src/cosmo.c:206-236)."""
    hp, c = nu_cosmo(tmp_path, Nmesh=16, Nsample=16)
    om = hp.Omega
    a_eval = np.array([0.1, 0.4, 1.0])

    def rhs(x, y, k):
        a = np.exp(x)
        H = np.sqrt(om / a ** 3 + 1 - om)
        dH = -1.5 * om / a ** 4 / H
        k2 = (k * 2997.92458) ** 2
        mu = 1 + k2 / (3 * (k2 + a * a * c.mass2(a)))
        return [y[1], -(2 + a * dH / H) * y[1] + 1.5 * om * mu * c.munu_1lpt(a, k) / (a ** 3 * H * H) * y[0]]

    for k in (0.05, 0.5):
        a0 = 1 / (1 + 200.0)
        s = integrate.solve_ivp(rhs, (np.log(a0), 0.0), [1, 1], t_eval=np.log(a_eval), args=(k,), method="DOP853",
                                rtol=1e-10, atol=1e-12)
        d = s.y[0] / s.y[0][-1]
        for i, a in enumerate(a_eval):
            assert c.growth_sd(0, k, a) == pytest.approx(d[i], rel=3e-4), (k, a)
    c.close()


def test_ascii_snapshot_layout(tmp_path):
    """The ASCII snapshot of builds without GADGET_STYLE (src/main.c:999-1027): particle
    count on the first line, then x y z vx vy vz with %12.6f, positions scaled to Mpc/h."""
    hp, c = cosmo_for(tmp_path, model="FOFR", Nsample=4, Nmesh=8)
    c.close()
    L = M.hlib()
    L.mgp_write_ascii.argtypes = [C.POINTER(M.HostParams), C.c_double, C.c_double, C.c_void_p, C.c_void_p, C.c_size_t]
    npart = 50
    rng = np.random.default_rng(2)
    pos = rng.uniform(0, hp.Box, size=(3, npart)).astype(np.float32)
    vel = (rng.normal(size=(3, npart)) * 300).astype(np.float32)
    assert L.mgp_write_ascii(C.byref(hp), 0.25, 3.0, pos.ctypes.data, vel.ctypes.data, npart) == 0
    lines = open(tmp_path / f"{hp.FileBase.decode()}_z3p000.0").read().splitlines()
    assert lines[0] == str(npart) and len(lines) == npart + 1
    assert len(lines[1]) == 6 * 12 + 5 and lines[1][12] == " "
    rows = np.array([[float(x) for x in l.split()] for l in lines[1:]])
    lf = hp.UnitLength_in_cm / 3.085678e24
    np.testing.assert_allclose(rows[:, :3], (pos.T * lf).astype(np.float32), atol=5e-7)
    np.testing.assert_allclose(rows[:, 3:], vel.T, atol=5e-7)
