"""GPU parity of the PM step kernels against the CPU oracle (FP64 restatement).

Each test runs one reference routine through the C ABI (libmgpu.so) and the
oracle on the same seeded inputs and compares within the tolerances the
kernels' FP32 arithmetic allows:
  density mesh          1e-6 relative to the cell mass 1 + delta
  spectra / forces      2e-5 of the field's maximum
  P(k)                  0.1 % per bin
  Disp, Kick, Drift     FP32 rounding of the same double expressions
"""
import numpy as np
import pytest

import mgpicola as M
import oracle as O
from helpers import lattice_particles, rel_err

pytestmark = pytest.mark.gpu

N, NS, BOX, OM = 32, 16, 64.0, 0.3
NBINS = 24


@pytest.fixture(scope="module")
def pos():
    return lattice_particles(NS, BOX, amp=1.5, seed=3)


def oracle_density(c, pos):
    d = O.deposit(c, pos)
    dx = O.real_view(d, c.nmesh).copy()
    dxfull = d.copy()
    O.fft_r2c_3d(c.nmesh, d)
    return d, dx, dxfull


def test_device_visible():
    info = M.device_info()
    print("device:", info, flush=True)
    assert info["cu_count"] > 0


def test_create_over_the_device_memory_fails_cleanly():
    """2048^3 particles (515 GB of particle arrays alone) are refused before any allocation, with the plan
    by part in the message; a context of normal size still works afterwards."""
    parts, total = M.memory_plan(2048, 2048, scale_dependent=1)
    assert total > M.device_info()["global_mem"]
    with pytest.raises(M.MGPUError) as e:
        M.MGPU(2048, 2048, 4096.0, 0.3, scale_dependent=1)
    msg = str(e.value)
    print(msg, flush=True)
    assert "(-2)" in msg or "-2" in msg
    assert "particles" in msg and "GB are free" in msg
    g = M.MGPU(32, 16, 64.0, 0.3)
    g.set_particles(lattice_particles(16, 64.0, amp=1.0, seed=1))
    g.pm_deposit()
    assert np.isfinite(g.get_mesh(M.MESH_DENSITY_X)).all()
    g.close()


@pytest.mark.parametrize("n,ns", [(32, 16), (24, 24), (48, 16)])
def test_deposit_matches_oracle(n, ns):
    p = lattice_particles(ns, BOX, amp=1.2, seed=5)
    g = M.MGPU(n, ns, BOX, OM)
    g.set_particles(p)
    g.pm_deposit()
    dx = g.get_mesh(M.MESH_DENSITY_X)
    c = O.cfg(n, ns, BOX, OM)
    ref = O.real_view(O.deposit(c, p), n)
    err = np.abs(dx - ref) / np.maximum(1.0, np.abs(1.0 + ref))
    assert err.max() < 1e-6, err.max()
    # mass conservation: sum delta = 0 up to rounding
    assert abs(dx.sum(dtype=np.float64)) < 1e-6 * n ** 3


def clustered_particles(n, ns, box, seed):
    """Half the particles in a tight clump straddling a corner shared by 8
    tiles, some exactly on the box edges, the rest uniform."""
    rng = np.random.default_rng(seed)
    npart = ns ** 3
    t = 16 if n % 16 == 0 else 8 if n % 8 == 0 else 4
    h = box / n
    p = rng.uniform(0.0, box, size=(3, npart))
    k = npart // 2
    p[:, :k] = t * h + rng.normal(0.0, 0.7 * h, size=(3, k))
    p[:, k:k + 8] = 0.0
    p[0, k + 8:k + 16] = np.nextafter(np.float32(box), np.float32(0))
    p = np.mod(p, box).astype(np.float32)
    p[p >= np.float32(box)] = np.nextafter(np.float32(box), np.float32(0))
    return p


@pytest.mark.parametrize("n,ns", [(12, 6), (16, 8), (24, 12), (32, 16)])
def test_deposit_clustered_matches_oracle(n, ns):
    """Tiles of 4, 8 and 16 cells (one, two or three per side), a clump over a
    tile corner and particles on the box edges: the bins' face classes and the
    periodic lower neighbours."""
    p = clustered_particles(n, ns, BOX, seed=n)
    print("n, ns =", n, ns, flush=True)
    g = M.MGPU(n, ns, BOX, OM)
    g.set_particles(p)
    g.pm_deposit()
    dx = g.get_mesh(M.MESH_DENSITY_X)
    c = O.cfg(n, ns, BOX, OM)
    ref = O.real_view(O.deposit(c, p), n)
    err = np.abs(dx - ref) / np.maximum(1.0, np.abs(1.0 + ref))
    assert err.max() < 1e-6, err.max()


@pytest.mark.parametrize("n,ns", [(12, 6), (24, 12), (32, 16)])
def test_tiled_gather_equals_direct_gather(n, ns):
    """The gather by the deposit's tiles gives the same Disp and the same mean
    (bit for bit) as the direct gather in particle order."""
    p = clustered_particles(n, ns, BOX, seed=n + 1)
    print("n, ns =", n, ns, flush=True)
    g = M.MGPU(n, ns, BOX, OM)
    g.set_particles(p)
    g.pm_density()
    g.forces()
    s_tiled = g.mesh_to_particles()
    d_tiled = g.get_disp()
    g.set_particles(p)  # new positions: the bins no longer apply, the direct gather runs
    s_direct = g.mesh_to_particles()
    d_direct = g.get_disp()
    assert np.array_equal(d_tiled, d_direct)
    assert np.array_equal(s_tiled, s_direct)
    c = O.cfg(n, ns, BOX, OM)
    d, _, _ = oracle_density(c, p)
    rdisp, _ = O.gather(c, p, O.forces(c, d, None))
    assert rel_err(d_tiled, rdisp) < 2e-5


@pytest.mark.parametrize("axis", [1, 2])
@pytest.mark.parametrize("mode", ["plain", "sd", "lpt"])
def test_rsd_multipoles_match_oracle(axis, mode):
    """PtoMesh_RSD + r2c + bin_up_RSD_power_spectrum against the oracle for the
    y and z lines of sight and each form of the velocity (no COLA term, the
    scale-dependent dD/dy field, D dDdy + D2 dD2dy): identical mode counts and
    bin k, P0 within 0.1 %, P2 and P4 within 0.1 % of P0."""
    n, ns, box, nbins = 32, 16, 64.0, 16
    p = clustered_particles(n, ns, box, seed=3 * axis + len(mode))
    rng = np.random.default_rng(11 * axis + len(mode))
    vel, l0, l1 = (rng.normal(0.0, sd, size=(3, ns ** 3)).astype(np.float32) for sd in (3.0, 1.0, 0.5))
    sd = 1 if mode == "sd" else 0
    cola = 0 if mode == "plain" else 1
    vnorm, s1, s2 = 0.37, 0.8, -0.2
    print("axis, mode =", axis, mode, flush=True)
    g = M.MGPU(n, ns, box, OM, scale_dependent=sd, use_cola=cola, pofk_nbins=nbins)
    g.set_particles(p, vel)
    g.set_lpt(0, l0)
    g.set_lpt(1, l1)
    P0, P2, P4, k, nm = g.rsd_power_spectrum(axis, vnorm, s1, s2)
    c = O.cfg(n, ns, box, OM)
    d = O.deposit_rsd(c, p, vel, l0, l1, axis, vnorm, sd, cola, s1, s2)
    O.fft_r2c_3d(n, d)
    nb, bt, sub, kmin, kmax = O.adjust_pofk(n, nbins, 0, 1, 0.0, 0.0)
    r0, r2, r4, rk, rn = O.rsd_multipoles(c, d, nb, bt, kmin, kmax, sub)
    assert np.array_equal(nm, rn)
    np.testing.assert_allclose(k, rk, rtol=1e-12)
    m = rn > 0
    scale = np.abs(r0[m])
    np.testing.assert_allclose(P0[m], r0[m], rtol=1e-3, atol=1e-6 * scale.max())
    for got, ref in ((P2, r2), (P4, r4)):
        assert np.all(np.abs(got[m] - ref[m]) <= 1e-3 * scale + 1e-6 * scale.max())
    # the velocities matter: the z and y multipoles of a clumped field differ from zero-velocity ones
    g.set_particles(p, np.zeros_like(vel))
    g.set_lpt(0, np.zeros_like(l0))
    g.set_lpt(1, np.zeros_like(l1))
    Q0, Q2, _, _, _ = g.rsd_power_spectrum(axis, vnorm, s1, s2)
    assert np.abs(Q2[m] - P2[m]).max() > 1e-3 * scale.max()


def test_neutrino_term_closed_form():
    """PtoMesh's neutrino term (src/auxPM.c:383-420): delta <- cdmfac delta +
    fac(k) cdelta_IC on every stored mode but k = 0, with fac quadratic in ln k
    (which the cubic Hermite slice reproduces exactly), compared with the closed
    form, both inside mgpu_pm_density and applied on its own."""
    n, ns, box = 32, 16, 64.0
    g = M.MGPU(n, ns, box, OM, scale_dependent=1)
    g.set_particles(lattice_particles(ns, box, amp=1.0, seed=9))
    g.pm_density()
    dk0 = g.get_mesh(M.MESH_DENSITY_K)
    rng = np.random.default_rng(4)
    shape = dk0.shape
    g.set_mesh(M.MESH_CDELTA1, rng.normal(size=shape) + 1j * rng.normal(size=shape))
    cd = g.get_mesh(M.MESH_CDELTA1)  # the FP32 values on the device
    kf = 2 * np.pi / box
    nk, y0 = 400, np.log(kf * 0.5)
    dy = (np.log(kf * n) - y0) / (nk - 1)
    yc, c0, c1, c2 = np.log(kf * 4), 0.7, -0.3, 0.05
    y = y0 + dy * np.arange(nk)
    tabs = [c0 + c1 * (y - yc) + c2 * (y - yc) ** 2, c1 + 2 * c2 * (y - yc), np.ones(nk), np.zeros(nk)]
    fac = M.KSlice(y0, dy, nk, *[t.ctypes.data_as(M.C.POINTER(M.C.c_double)) for t in tabs], 1.0)
    f = np.fft.fftfreq(n) * n
    n2 = f[:, None, None] ** 2 + f[None, :, None] ** 2 + np.arange(shape[2])[None, None, :] ** 2
    n2[0, 0, 0] = 1.0
    yk = np.log(kf * np.sqrt(n2))
    fk = c0 + c1 * (yk - yc) + c2 * (yk - yc) ** 2
    fk[0, 0, 0] = 0.0
    cdmfac = 0.953

    def term(d):
        r = cdmfac * d + fk * cd
        r[0, 0, 0] = d[0, 0, 0]
        return r

    g.set_neutrinos(cdmfac, fac)
    g.pm_density()
    got = g.get_mesh(M.MESH_DENSITY_K)
    assert rel_err(got, term(dk0)) < 1e-6
    assert got[0, 0, 0] == dk0[0, 0, 0]
    g.neutrino_response()
    assert rel_err(g.get_mesh(M.MESH_DENSITY_K), term(got)) < 1e-6
    g.set_neutrinos(1.0, None)
    g.pm_density()
    assert rel_err(g.get_mesh(M.MESH_DENSITY_K), dk0) == 0.0
    with pytest.raises(M.MGPUError):
        g.neutrino_response()
    g.close()


def test_deposit_is_deterministic(pos):
    g = M.MGPU(N, NS, BOX, OM)
    g.set_particles(pos)
    g.pm_deposit()
    a = g.get_mesh(M.MESH_DENSITY_X)
    g.pm_deposit()
    b = g.get_mesh(M.MESH_DENSITY_X)
    assert np.array_equal(a, b)


@pytest.mark.parametrize("n", [8, 12, 16, 24, 32, 48, 64, 96])
def test_r2c_matches_oracle_and_numpy(n):
    ns = n // 2
    p = lattice_particles(ns, BOX, amp=1.0, seed=n)
    g = M.MGPU(n, ns, BOX, OM)
    g.set_particles(p)
    g.pm_deposit()
    dx = g.get_mesh(M.MESH_DENSITY_X)
    g.pm_transform()
    dk = g.get_mesh(M.MESH_DENSITY_K)
    ref = np.fft.rfftn(dx.astype(np.float64))
    assert rel_err(dk, ref) < 2e-6


def check_pofk(got, ref):
    """Identical mode counts, mean |k| to FP64 rounding and the binned power
    (no shot-noise constant) within 0.1 % in every bin with k > 0."""
    pk, kk, nm = got
    rp, rk, rn = ref
    assert np.array_equal(nm, rn)
    m = (rn > 0) & (rk > 0)  # the k = 0 bin holds the mean, zero by construction
    np.testing.assert_allclose(kk[m], rk[m], rtol=1e-10)  # FP64 sums in another order
    np.testing.assert_allclose(pk[m], rp[m], rtol=1e-3, atol=0)


def test_pofk_matches_oracle(pos):
    g = M.MGPU(N, NS, BOX, OM, pofk_nbins=NBINS, pofk_subtract_shotnoise=0)
    g.set_particles(pos)
    got = g.pm_density(pofk=True)
    c = O.cfg(N, NS, BOX, OM)
    d, _, _ = oracle_density(c, pos)
    check_pofk(got, O.pofk(c, d, NBINS, 0, 0.0, float(N), 0))


@pytest.mark.parametrize("nbins", [2000, 20000, 21000])
def test_pofk_many_bins_matches_oracle(pos, nbins):
    """Bin counts beyond the LDS of the fused x pass fall back to k_pofk over
    bin ranges: 20000 bins fit the 8-column density pair of the LCDM step but
    not the 16-column x pass of the density call, 21000 fit neither; every path
    agrees with the oracle."""
    c = O.cfg(N, NS, BOX, OM)
    d, _, _ = oracle_density(c, pos)
    ref = O.pofk(c, d, nbins, 0, 0.0, float(N), 0)
    g = M.MGPU(N, NS, BOX, OM, pofk_nbins=nbins, pofk_subtract_shotnoise=0)
    g.set_particles(pos)
    check_pofk(g.pm_density(pofk=True), ref)   # first call: k_pofk, records the geometry
    check_pofk(g.pm_density(pofk=True), ref)   # fused x pass where the bins fit
    _, got = g.get_displacements(pofk=True)
    check_pofk(got, ref)
    g.close()


@pytest.mark.parametrize("bintype,kmin,kmax", [(0, 0.0, float(N)), (1, 1.0, float(N))])
def test_pofk_binned_in_transform_matches_oracle(pos, bintype, kmin, kmax):
    """From the second call on the power is binned inside the r2c's last pass;
    mode counts and mean |k| come from the first call's full binning."""
    g = M.MGPU(N, NS, BOX, OM, pofk_nbins=NBINS, pofk_bintype=bintype, pofk_kmin=kmin, pofk_kmax=kmax,
               pofk_subtract_shotnoise=0)
    g.set_particles(pos)
    first = g.pm_density(pofk=True)
    second = g.pm_density(pofk=True)
    c = O.cfg(N, NS, BOX, OM)
    d, _, _ = oracle_density(c, pos)
    ref = O.pofk(c, d, NBINS, bintype, kmin, kmax, 0)
    for got in (first, second):
        check_pofk(got, ref)


def test_forces_and_gather_lcdm(pos):
    g = M.MGPU(N, NS, BOX, OM)
    g.set_particles(pos)
    g.pm_density()
    g.forces()
    c = O.cfg(N, NS, BOX, OM)
    d, _, _ = oracle_density(c, pos)
    f = O.forces(c, d, None)
    for a, mid in enumerate((M.MESH_FORCE_X, M.MESH_FORCE_Y, M.MESH_FORCE_Z)):
        assert rel_err(g.get_mesh(mid), O.real_view(f[a], N)) < 2e-5
    sumD = g.mesh_to_particles()
    disp = g.get_disp()
    rdisp, rsum = O.gather(c, pos, f)
    assert rel_err(disp, rdisp) < 2e-5
    assert np.abs(sumD - rsum).max() < 2e-5 * np.abs(rdisp).max()


MG_CASES = {
    "fofr_screen": dict(model=O.FOFR, scr=1, mg=dict(a=0.5, coupling=1.0 / 3.0, mass2=2.0e4, phi_crit=3e-8)),
    "fofr_noscreen": dict(model=O.FOFR, scr=0, mg=dict(a=0.5, coupling=1.0 / 3.0, mass2=2.0e4, phi_crit=3e-8)),
    "dgp_screen": dict(model=O.DGP, scr=1, filt=O.FILTER_GAUSSIAN,
                       mg=dict(a=0.7, coupling=0.12, dgp_fac=0.05, rsmooth=2.0)),
    "dgp_noscreen": dict(model=O.DGP, scr=0, mg=dict(a=0.7, coupling=0.12)),
    "mbeta_screen": dict(model=O.MBETA, scr=1, mg=dict(a=0.8, coupling=0.5, mass2=5.0e3, phi_crit=2e-8)),
    "jbd": dict(model=O.BRANSDICKE, scr=0, mg=dict(a=0.6, geff=1.04)),
}


@pytest.mark.parametrize("case", sorted(MG_CASES))
def test_fifth_force_matches_oracle(pos, case):
    cs = MG_CASES[case]
    filt = cs.get("filt", O.FILTER_NONE)
    mgd = dict(a=1.0, coupling=0.0, mass2=0.0, phi_crit=0.0, dgp_fac=0.0, rsmooth=1.0, geff=1.0)
    mgd.update(cs["mg"])
    g = M.MGPU(N, NS, BOX, OM, model=cs["model"], mg_active=1, include_screening=cs["scr"], filter=filt)
    g.set_particles(pos)
    g.pm_density()
    g.fifth_force(M.MG(**mgd))
    c = O.cfg(N, NS, BOX, OM, model=cs["model"], mg_active=1, include_screening=cs["scr"], filter=filt)
    d, _, dxfull = oracle_density(c, pos)
    phik = O.fifth_force(c, O.MG(**mgd), d, dxfull)
    if phik is not None:
        assert rel_err(g.get_mesh(M.MESH_PHI_K), O.complex_view(phik, N)) < 3e-5
    g.forces()
    f = O.forces(c, d, phik)
    for a, mid in enumerate((M.MESH_FORCE_X, M.MESH_FORCE_Y, M.MESH_FORCE_Z)):
        assert rel_err(g.get_mesh(mid), O.real_view(f[a], N)) < 3e-5


@pytest.mark.parametrize("case", ["lcdm", "fofr_screen", "dgp_screen", "mbeta_screen"])
@pytest.mark.parametrize("n", [32, 48])
def test_fused_step_equals_separate_calls(pos, case, n):
    """GetDisplacements with paired x passes (phi(k) never stored) against
    PtoMesh, ComputeFifthForce, Forces and MtoParticles called one by one:
    force meshes, Disp, the mean and P(k).  The first density call of each
    context takes the unpaired path (it records the P(k) mode counts), so the
    paired path runs from the second call on."""
    ns = n // 2
    p = lattice_particles(ns, BOX, amp=1.5, seed=n)
    if case == "lcdm":
        kw, mg = dict(), None
    else:
        cs = MG_CASES[case]
        kw = dict(model=cs["model"], mg_active=1, include_screening=cs["scr"], filter=cs.get("filt", O.FILTER_NONE))
        mgd = dict(a=1.0, coupling=0.0, mass2=0.0, phi_crit=0.0, dgp_fac=0.0, rsmooth=1.0, geff=1.0)
        mgd.update(cs["mg"])
        mg = M.MG(**mgd)
    print("case, n =", case, n, flush=True)
    g = M.MGPU(n, ns, BOX, OM, pofk_nbins=n // 2, **kw)
    g.set_particles(p)
    g.get_displacements(mg=mg, pofk=True)  # unpaired: records the mode counts
    s_fused, (pk_f, k_f, n_f) = g.get_displacements(mg=mg, pofk=True)
    f_fused = [g.get_mesh(m) for m in (M.MESH_FORCE_X, M.MESH_FORCE_Y, M.MESH_FORCE_Z)]
    d_fused = g.get_disp()
    dk_fused = g.get_mesh(M.MESH_DENSITY_K)
    pk_s, k_s, n_s = g.pm_density(pofk=True)
    dk_sep = g.get_mesh(M.MESH_DENSITY_K)
    if mg is not None:
        g.fifth_force(mg)
    g.forces()
    s_sep = g.mesh_to_particles()
    f_sep = [g.get_mesh(m) for m in (M.MESH_FORCE_X, M.MESH_FORCE_Y, M.MESH_FORCE_Z)]
    d_sep = g.get_disp()
    assert rel_err(dk_fused, dk_sep) < 1e-6
    for a in range(3):
        assert rel_err(f_fused[a], f_sep[a]) < 1e-6, a
    assert rel_err(d_fused, d_sep) < 1e-6
    assert np.abs(s_fused - s_sep).max() <= 1e-6 * np.abs(d_sep).max()
    assert np.array_equal(n_f, n_s) and np.array_equal(k_f, k_s)
    np.testing.assert_allclose(pk_f, pk_s, rtol=1e-5, atol=1e-9 * np.abs(pk_s).max())


def test_fofr_screening_is_active(pos):
    """The f(R) case above must actually screen some cells (else it pins nothing)."""
    c = O.cfg(N, NS, BOX, OM, model=O.FOFR, mg_active=1, include_screening=1)
    d, _, dxfull = oracle_density(c, pos)
    mg = O.MG(a=0.5, coupling=1.0 / 3.0, mass2=2.0e4, phi_crit=3e-8, dgp_fac=0, rsmooth=1, geff=1)
    p_scr = O.fifth_force(c, mg, d.copy(), dxfull)
    c0 = O.cfg(N, NS, BOX, OM, model=O.FOFR, mg_active=1, include_screening=0)
    p_unscr = O.fifth_force(c0, mg, d.copy(), dxfull)
    diff = rel_err(O.complex_view(p_scr, N), O.complex_view(p_unscr, N))
    assert diff > 1e-2


@pytest.mark.parametrize("sd", [0, 1])
def test_kick_drift_match_oracle(pos, sd):
    rng = np.random.default_rng(11)
    vel = (rng.normal(size=(3, NS ** 3)) * 3.0).astype(np.float32)
    l0 = (rng.normal(size=(3, NS ** 3)) * 0.7).astype(np.float32)
    l1 = (rng.normal(size=(3, NS ** 3)) * 0.1).astype(np.float32)
    g = M.MGPU(N, NS, BOX, OM, scale_dependent=sd)
    g.set_particles(pos, vel)
    g.set_lpt(0, l0)
    g.set_lpt(1, l1)
    g.pm_density()
    g.forces()
    sumD = g.mesh_to_particles()
    disp = g.get_disp()
    dda, A, s1, s2 = 0.37, 0.4, 0.8, -0.2
    sV = g.kick(dda, A, s1, s2, sumD)
    rvel, rdisp = vel.copy(), disp.copy()
    if sd:  # kick uses slot0 alone, drift slot1 alone (each already the order-1 + order-2 sum)
        rsV = O.kick(rvel, rdisp, l0, np.zeros_like(l0), 0, 0, 1, OM, dda, A, 1, sumD)
    else:
        rsV = O.kick(rvel, rdisp, l0, l1, s1, s2, 0, OM, dda, A, 1, sumD)
    gpos, gvel = g.get_particles()
    np.testing.assert_array_equal(gvel, rvel)
    np.testing.assert_array_equal(g.get_disp(), rdisp)
    np.testing.assert_allclose(sV, rsV, rtol=1e-12, atol=1e-12)
    dyyy, t1, t2 = 0.9, 0.05, 0.01
    g.drift(dyyy, t1, t2, sV)
    rpos = pos.copy()
    if sd:
        O.drift(rpos, rvel, l1, np.zeros_like(l1), 0, 0, 1, BOX, dyyy, 1, rsV)
    else:
        O.drift(rpos, rvel, l0, l1, t1, t2, 0, BOX, dyyy, 1, rsV)
    gpos, _ = g.get_particles()
    np.testing.assert_array_equal(gpos, rpos)


@pytest.mark.parametrize("sd", [0, 1])
def test_fused_kick_drift_equals_separate_passes(pos, sd):
    """mgpu_kick_drift against mgpu_kick + mgpu_drift on the same state: the
    velocities and the returned mean velocity bit for bit; the positions within
    FP32 rounding (the drift takes the mean velocity from before the kick, equal
    to the mean after it up to rounding because the kick term has zero mean)."""
    rng = np.random.default_rng(12)
    vel = (rng.normal(size=(3, NS ** 3)) * 3.0).astype(np.float32)
    l0 = (rng.normal(size=(3, NS ** 3)) * 0.7).astype(np.float32)
    l1 = (rng.normal(size=(3, NS ** 3)) * 0.1).astype(np.float32)
    l0 -= l0.mean(axis=1, keepdims=True)  # LPT fields have zero mean
    l1 -= l1.mean(axis=1, keepdims=True)
    dda, A, s1k, s2k, dyyy, s1d, s2d = 0.37, 0.4, 0.8, -0.2, 0.9, 0.05, 0.01

    def state():
        g = M.MGPU(N, NS, BOX, OM, scale_dependent=sd)
        g.set_particles(pos, vel)
        g.set_lpt(0, l0)
        g.set_lpt(1, l1)
        g.pm_density()
        g.forces()
        return g, g.mesh_to_particles()

    g, sumD = state()
    sV0 = vel.astype(np.float64).mean(axis=1)
    sV = g.kick(dda, A, s1k, s2k, sumD)
    g.drift(dyyy, s1d, s2d, sV)
    p_sep, v_sep = g.get_particles()
    g.close()
    g, sumD2 = state()
    np.testing.assert_array_equal(sumD2, sumD)
    sV_f = g.kick_drift(dda, A, s1k, s2k, sumD, dyyy, s1d, s2d, sV0)
    p_f, v_f = g.get_particles()
    np.testing.assert_array_equal(v_f, v_sep)
    np.testing.assert_array_equal(sV_f, sV)
    d = np.abs(p_f - p_sep)
    d = np.minimum(d, BOX - d)  # periodic
    assert d.max() < 1e-5 * BOX, d.max()
    assert np.abs(sV - sV0).max() < 1e-3 * np.abs(sV0).max()  # the premise: the kick keeps the mean
    with pytest.raises(M.MGPUError):  # Disp was not reduced by sumD: a further kick needs a new gather
        g.kick(dda, A, s1k, s2k, [0.0, 0.0, 0.0])
    g.close()


@pytest.mark.parametrize("n,ns,sd", [(32, 16, 0), (48, 16, 1), (64, 32, 1)])
def test_fft_tile_width_changes_no_result(n, ns, sd):
    """Tiles of 8 and 16 columns give bit-identical density, forces and Disp; the
    autotuner times both (where 16 fits) and keeps one of them."""
    p = clustered_particles(n, ns, BOX, seed=n)
    out = []
    for t in (8, 16):
        g = M.MGPU(n, ns, BOX, OM, scale_dependent=sd)
        g.set_fft_tile(0, t)
        assert g.fft_tile(0) == t
        g.set_particles(p)
        g.pm_density()
        g.forces()
        g.mesh_to_particles()
        out.append([g.get_mesh(M.MESH_DENSITY_K)] + [g.get_mesh(q) for q in (M.MESH_FORCE_X, M.MESH_FORCE_Y,
                                                                             M.MESH_FORCE_Z)] + [g.get_disp()])
        g.close()
    for a, b in zip(*out):
        assert np.array_equal(a, b)
    g = M.MGPU(n, ns, BOX, OM, scale_dependent=sd)
    ms = g.autotune_fft()
    assert ms[0] > 0 and ms[1] > 0  # both widths fit at these sizes
    assert g.fft_tile(0) == (16 if ms[1] <= ms[0] else 8)
    if sd:
        assert ms[2] > 0 and g.fft_tile(1) in (8, 16)
    with pytest.raises(M.MGPUError):
        g.set_fft_tile(0, 12)
    g.close()
