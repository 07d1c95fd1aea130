"""End-to-end GPU runs through the C host (libmgpicola -> libmgpu) against the
CPU oracle, and the initial conditions against the oracle's 2LPT restatement.

Tolerances: positions after the final step within 1e-4 of the box (FP32 GPU vs
FP64 oracle); initial displacements within 1e-5 of their maximum.
"""
import os

import numpy as np
import pytest

import mgpicola as M
import oracle as O
import workloads as W
from e2e import oracle_run, periodic_diff
from helpers import rel_err

pytestmark = pytest.mark.gpu


def make_params(tmp_path, name, nsteps, model=None, **over):
    wmodel, vals, _ = W.values_for(name, str(tmp_path), **over)
    model = model or wmodel
    with open(vals["OutputRedshiftFile"], "w") as f:
        f.write(f"0, {nsteps}\n")
    pf = os.path.join(str(tmp_path), "params.txt")
    M.write_paramfile(pf, vals, model=model)
    return M.read_params(pf, model=model)


@pytest.mark.parametrize("case", [
    dict(name="lcdm_64_128", nsteps=10, over=dict(Nmesh=64, Nsample=32, Box=128.0)),
    dict(name="lcdm_64_128", nsteps=10, over=dict()),
    dict(name="lcdm_64_128", nsteps=5, over=dict(Nmesh=48, Nsample=24, Box=96.0, modified_gravity_active=1,
                                                 include_screening=1)),
    # f(R) with scale-dependent growth and screening (the benchmark's path) at 32^3 / 64^3
    dict(name="fofr_sd_512_1024", nsteps=6, over=dict(Nmesh=64, Nsample=32, Box=64.0, pofk_nbins=32)),
])
def test_end_to_end_matches_oracle(tmp_path, case):
    hp = make_params(tmp_path, case["name"], case["nsteps"], **case["over"])
    print("case:", case, flush=True)
    sim = M.Sim(hp)
    sim.ic()
    while not sim.finished:
        sim.advance()
    gpos, gvel = sim.gpu.get_particles()
    sim.close()
    cosmo = M.Cosmo(hp)
    rpos, rvel, _ = oracle_run(hp, cosmo, case["nsteps"])
    d = np.abs(periodic_diff(gpos, rpos, hp.Box))
    print("max |dpos| / Box =", d.max() / hp.Box, flush=True)
    assert d.max() <= 1e-4 * hp.Box
    # output files of the reference's formats were written
    files = os.listdir(str(tmp_path))
    assert any(f.startswith(case["name"] + "_z0p000") for f in files), files
    assert any(f.startswith("info_" + case["name"]) for f in files), files
    if hp.pofk_compute_every_step:
        assert sum(f.startswith("pofk_" + case["name"]) for f in files) == case["nsteps"] + 1, files


def run_gpu(hp, track=False):
    sim = M.Sim(hp)
    if track:
        sim.set_learned(None, track=True)  # logs the P(k) of every force evaluation; no multiplier
    sim.ic()
    while not sim.finished:
        sim.advance()
    gpos, gvel = sim.gpu.get_particles()
    log = sim.log() if track else None
    sim.close()
    return gpos, gvel, log


@pytest.mark.parametrize("case", [
    # nDGP rc H0 = 1 with density screening (Gaussian smoothing)
    dict(name="dgp_96_192", model="DGP", nsteps=6, over=dict(Nmesh=64, Nsample=32, Box=64.0)),
    # Jordan-Brans-Dicke (w = 100): G_eff(a) from the JBD background
    dict(name="lcdm_64_128", model="BRANSDICKE", nsteps=6,
         over=dict(Nmesh=64, Nsample=32, Box=64.0, modified_gravity_active=1)),
    # m(a)-beta(a) symmetron-like model with screening and scale-dependent growth
    dict(name="lcdm_64_128", model="MBETA", nsteps=6,
         over=dict(Nmesh=64, Nsample=32, Box=64.0, modified_gravity_active=1, include_screening=1)),
    # the benchmark's path (f(R) f_R0 = 1e-5, scale-dependent, screening, P(k) every step) at 128^3 / 256^3:
    # two-tile column passes, lattice fold m = 2 from a 256 mesh
    dict(name="fofr_sd_512_1024", model=None, nsteps=4, over=dict(Nmesh=256, Nsample=128, Box=256.0,
                                                                  pofk_nbins=128)),
])
def test_end_to_end_models_match_oracle(tmp_path, case):
    """Whole runs of the other gravity models and of the benchmark's path at
    128^3 particles on a 256^3 mesh: final positions within 1e-4 of the box."""
    hp = make_params(tmp_path, case["name"], case["nsteps"], model=case["model"], **case["over"])
    print("case:", case, flush=True)
    gpos, gvel, _ = run_gpu(hp)
    cosmo = M.Cosmo(hp)
    rpos, rvel, _ = oracle_run(hp, cosmo, case["nsteps"])
    d = np.abs(periodic_diff(gpos, rpos, hp.Box))
    v = np.abs(gvel.astype(np.float64) - rvel)
    print("max |dpos| / Box =", d.max() / hp.Box, " max |dvel| / max |vel| =", v.max() / np.abs(rvel).max(),
          flush=True)
    assert d.max() <= 1e-4 * hp.Box
    assert v.max() <= 1e-3 * np.abs(rvel).max()


def test_fused_kick_drift_run_matches_separate_passes(tmp_path):
    """The whole 10-step LCDM workload with Kick and Drift in one pass (the
    default: the drift subtracts the mean velocity from before the kick) and
    with the reference's separate passes (fuse_kick_drift = 0: the mean of the
    kicked velocities, src/main.c:771): final positions agree to 1e-5 of the
    box, and both agree with the oracle (separate passes) to 1e-4."""
    nsteps = 10
    out = []
    for fuse in (1, 0):
        hp = make_params(tmp_path, "lcdm_64_128", nsteps)
        hp.fuse_kick_drift = fuse
        gpos, gvel, _ = run_gpu(hp)
        out.append((gpos, gvel))
    cosmo = M.Cosmo(hp)
    rpos, rvel, _ = oracle_run(hp, cosmo, nsteps)
    d_fs = np.abs(periodic_diff(out[0][0], out[1][0], hp.Box)).max() / hp.Box
    d_fo = np.abs(periodic_diff(out[0][0], rpos, hp.Box)).max() / hp.Box
    d_so = np.abs(periodic_diff(out[1][0], rpos, hp.Box)).max() / hp.Box
    print(f"max |dpos|/Box: fused vs separate {d_fs:.2e}, fused vs oracle {d_fo:.2e}, separate vs oracle {d_so:.2e}",
          flush=True)
    assert d_fs <= 1e-5
    assert d_fo <= 1e-4 and d_so <= 1e-4


def test_end_to_end_pofk_every_step_and_density(tmp_path):
    """The plain LCDM workload (64^3 particles, 128^3 mesh, Box 256, 10 steps)
    with P(k) at every force evaluation: each of the 11 P(k) (raw binned power,
    no shot-noise constant) within 0.1 % per bin of the oracle's on its own
    trajectory, identical mode counts; then the density of the final positions
    against the oracle's deposit of the same positions to 1e-6 of 1 + delta."""
    nsteps = 10
    hp = make_params(tmp_path, "lcdm_64_128", nsteps, pofk_compute_every_step=1, pofk_nbins=64,
                     pofk_subtract_shotnoise=0)
    gpos, gvel, log = run_gpu(hp, track=True)
    assert len(log) == nsteps + 1
    cosmo = M.Cosmo(hp)
    nb, bt, sub, kmin, kmax = O.adjust_pofk(hp.Nmesh, hp.pofk_nbins, hp.pofk_bintype, 0, 0.0, 0.0)
    rpos, rvel, _, rpk = oracle_run(hp, cosmo, nsteps, pofk=(nb, bt, sub, kmin, kmax))
    d = np.abs(periodic_diff(gpos, rpos, hp.Box))
    assert d.max() <= 1e-4 * hp.Box
    worst = 0.0
    for step, ((a, _, pk, kk, nm), (rp, rk, rn)) in enumerate(zip(log, rpk)):
        assert np.array_equal(nm, rn), step
        sel = (rn > 0) & (rk > 0)  # the k = 0 bin holds the mean, zero by construction
        np.testing.assert_allclose(kk[sel], rk[sel], rtol=1e-10)  # FP64 sums in another order
        e = np.abs(pk[sel] - rp[sel]) / np.abs(rp[sel])
        worst = max(worst, e.max())
        print(f"step {step} a {a:.4f}: P(k) max rel err {e.max():.2e}", flush=True)
        assert e.max() < 1e-3, step
    print("worst per-bin P(k) error over the run:", worst, flush=True)
    g = M.MGPU(hp.Nmesh, hp.Nsample, hp.Box, hp.Omega)
    g.set_particles(gpos)
    g.pm_deposit()
    dx = g.get_mesh(M.MESH_DENSITY_X)
    g.close()
    ref = O.real_view(O.deposit(O.cfg(hp.Nmesh, hp.Nsample, hp.Box, hp.Omega), gpos), hp.Nmesh)
    err = np.abs(dx - ref) / np.maximum(1.0, np.abs(1.0 + ref))
    assert err.max() < 1e-6, err.max()
    assert ref.max() > 10.0  # the z = 0 field is clustered


@pytest.mark.parametrize("n,ns,sd", [(32, 16, 0), (32, 32, 0), (24, 12, 1), (48, 24, 1), (24, 24, 1), (48, 16, 1),
                                     (32, 12, 1)])
def test_ic_matches_oracle(n, ns, sd):
    """ZA/2LPT displacements (non scale-dependent) and the stored spectra plus
    both-order LPT fields (scale-dependent) against the oracle; the last three
    cases cover the lattice fold with m = 1 and m = 3 and the full-mesh
    read-out (N not a multiple of Nsample)."""
    box = 2.0 * n
    rng = np.random.default_rng(n)
    kf = 2 * np.pi / box
    n2 = np.arange(3 * (n // 2) ** 2 + 1)
    k = kf * np.sqrt(n2)
    pk = np.where(n2 > 0, 2e4 * k / (1 + (k / 0.05) ** 2.5), 0.0)
    g = M.MGPU(n, ns, box, 0.3, scale_dependent=sd)
    g.ic_displacement_fields(5001, pk)
    c = O.cfg(n, ns, box, 0.3)
    cd = O.ic_delta(c, 5001, pk)
    assert rel_err(g.get_mesh(M.MESH_CDELTA1), O.complex_view(cd, n)) < 1e-6
    if sd:
        (cd2,) = O.ic_displacements(c, cd, 1)
        assert rel_err(g.get_mesh(M.MESH_CDELTA2), O.complex_view(cd2, n)) < 5e-6
        g1 = rng.uniform(0.5, 1.5, size=n2.size)
        g2 = rng.uniform(-1e-4, 1e-4, size=n2.size)
        g.lpt_fields(0, g1, g2)
        ref = O.lpt_field(c, cd, g1).astype(np.float64) + O.lpt_field(c, cd2, g2).astype(np.float64)
        assert rel_err(g.get_lpt(0), ref) < 1e-5
    else:
        za, lp = O.ic_displacements(c, cd, 0)
        assert rel_err(g.get_lpt(0), za) < 1e-5
        assert rel_err(g.get_lpt(1), lp) < 2e-5


@pytest.mark.parametrize("sphere,fixed,inverted", [(1, 0, 0), (0, 1, 0), (0, 0, 1), (1, 1, 1)])
@pytest.mark.parametrize("n,ns,sd", [(32, 16, 0), (24, 24, 0), (48, 24, 1)])
def test_ic_variants_match_oracle(n, ns, sd, sphere, fixed, inverted):
    """SphereMode = 1 (|k| < k_Nyquist of the particle grid), amplitude-fixed
    modes (|delta| = sqrt(P)) and inverted initial conditions (delta -> -delta),
    src/2LPT.c:347-412, against the oracle on the same RNG stream; with the
    amplitude fixed, every retained mode has |delta|^2 = P / Box^3 exactly up to
    FP32 rounding."""
    box = 2.0 * n
    kf = 2 * np.pi / box
    n2 = np.arange(3 * (n // 2) ** 2 + 1)
    k = kf * np.sqrt(n2)
    pk = np.where(n2 > 0, 2e4 * k / (1 + (k / 0.05) ** 2.5), 0.0)
    print("case", n, ns, sd, sphere, fixed, inverted, flush=True)
    g = M.MGPU(n, ns, box, 0.3, scale_dependent=sd)
    g.ic_displacement_fields(5001, pk, sphere_mode=sphere, amplitude_fixed=fixed, inverted=inverted)
    c = O.cfg(n, ns, box, 0.3)
    cd = O.ic_delta(c, 5001, pk, sphere, fixed, inverted)
    got = g.get_mesh(M.MESH_CDELTA1)
    ref = O.complex_view(cd, n)
    assert rel_err(got, ref) < 1e-6
    if fixed:
        kx = np.rint(np.fft.fftfreq(n) * n).astype(np.int64)
        m2 = kx[:, None, None] ** 2 + kx[None, :, None] ** 2 + np.arange(n // 2 + 1)[None, None, :] ** 2
        nz = np.abs(ref) > 0
        want = pk[m2] / box ** 3
        np.testing.assert_allclose(np.abs(got[nz]) ** 2, want[nz], rtol=2e-6)
    if sd:
        (cd2,) = O.ic_displacements(c, cd, 1)
        assert rel_err(g.get_mesh(M.MESH_CDELTA2), O.complex_view(cd2, n)) < 5e-6
    else:
        za, lp = O.ic_displacements(c, cd, 0)
        assert rel_err(g.get_lpt(0), za) < 1e-5
        assert rel_err(g.get_lpt(1), lp) < 2e-5
    g.close()


def test_lpt_slices_match_host_tables(tmp_path):
    """Growth evaluated on the GPU from ln k slices equals the host's per-|n|^2
    tables (both restate the bicubic lookup of src/cosmo.c:987-1010)."""
    hp = make_params(tmp_path, "lcdm_64_128", 4, Nmesh=32, Nsample=16, Box=64.0, modified_gravity_active=1,
                     include_screening=1)
    hp.scale_dependent = 1
    cosmo = M.Cosmo(hp)
    g = M.MGPU(32, 16, 64.0, 0.3, model=M.FOFR, mg_active=1, include_screening=1, scale_dependent=1)
    rng = np.random.default_rng(3)
    pk = np.where(np.arange(M.n2_table_len(32)) > 0, 1e4, 0.0)
    g.ic_displacement_fields(5001, pk)
    for field, A, AFF in ((M.FIELD_deltaD, 0.2, 0.25), (M.FIELD_ddDddy, 0.5, 0.55), (M.FIELD_D, 0.05, 0.05)):
        g1, g2 = cosmo.lpt_tables(field, A, AFF)
        g.lpt_fields(0, g1, g2)
        ref = g.get_lpt(0)
        o1, o2, buf = M.lpt_kslices(cosmo, field, A, AFF)
        M.lpt_fields_slices(g, 1, o1, o2)
        assert rel_err(g.get_lpt(1), ref) < 1e-6
    del rng


@pytest.mark.parametrize("n,ns", [(32, 16), (24, 24), (48, 16), (32, 12)])
def test_lpt_two_slot_fold_matches_single_slot(tmp_path, n, ns):
    """The two-slot call (one pass over the stored spectra for both LPT slots)
    equals two one-slot calls: lattice fold with m = 2, 1, 3 and the full-mesh
    path (N not a multiple of Nsample)."""
    hp = make_params(tmp_path, "lcdm_64_128", 4, Nmesh=n, Nsample=ns, Box=2.0 * n, modified_gravity_active=1,
                     include_screening=1)
    hp.scale_dependent = 1
    cosmo = M.Cosmo(hp)
    g = M.MGPU(n, ns, 2.0 * n, 0.3, model=M.FOFR, mg_active=1, include_screening=1, scale_dependent=1)
    k = np.arange(M.n2_table_len(n))
    g.ic_displacement_fields(5001, np.where(k > 0, 1e4 / (1.0 + 0.01 * k), 0.0))
    a1, a2, ba = M.lpt_kslices(cosmo, M.FIELD_ddDddy, 0.5, 0.55)
    b1, b2, bb = M.lpt_kslices(cosmo, M.FIELD_deltaD, 0.5, 0.55)
    M.lpt_fields_slices(g, 0, a1, a2)
    M.lpt_fields_slices(g, 1, b1, b2)
    r0, r1 = g.get_lpt(0), g.get_lpt(1)
    M.lpt_fields_slices2(g, (a1, a2), (b1, b2))
    assert rel_err(g.get_lpt(0), r0) < 1e-6
    assert rel_err(g.get_lpt(1), r1) < 1e-6
    assert rel_err(r0, r1) > 1e-3  # the two slots hold different fields
    del ba, bb


def test_rsd_files_written_every_step_and_at_output(tmp_path):
    """pofk_compute_rsd_pofk = 1 (scale-dependent f(R)): the RSD multipole file is
    written at every step but the first of each output interval and at the
    output, in the reference's format (7 columns, k > 0)."""
    hp = make_params(tmp_path, "fofr_sd_512_1024", 3, Nmesh=32, Nsample=16, Box=64.0, pofk_nbins=16,
                     pofk_compute_rsd_pofk=1)
    sim = M.Sim(hp)
    sim.ic()
    while not sim.finished:
        sim.advance()
    sim.close()
    files = sorted(f for f in os.listdir(str(tmp_path)) if f.startswith("pofk_RSD_"))
    # 3 steps from z = 19: steps 2 and 3 in the loop, and the output at z = 0
    assert len(files) == 3, files
    for f in files:
        rows = np.loadtxt(os.path.join(str(tmp_path), f), ndmin=2)
        assert rows.shape[1] == 7 and rows.shape[0] > 3
        assert np.all(np.isfinite(rows)) and np.all(rows[:, 0] > 0)
        assert np.all(rows[:, 4:] >= 0)


def test_fofrnu_run_writes_cdm_and_total_pofk(tmp_path):
    """f(R) with massive neutrinos (FOFRNU, the reference's nu0.2 CAMB tables at
    32^3 in a 512 Mpc/h box): the run writes the CDM and the total matter P(k)
    at every step (src/auxPM.c:370-371, :422-426).  Without shot-noise
    subtraction total/CDM lies between (1 - fnu)^2 and 1 and falls with k, as
    the neutrinos stream out of the small scales."""
    import test_host as H
    outs = tmp_path / "outs.dat"
    outs.write_text("0, 3\n")
    hp, c = H.nu_cosmo(tmp_path, Nmesh=32, Nsample=16, pofk_nbins=16, pofk_compute_rsd_pofk=0,
                       pofk_subtract_shotnoise=0, OutputRedshiftFile=str(outs))
    c.close()
    hp.growth_on_gpu = 1
    sim = M.Sim(hp)
    sim.ic()
    while not sim.finished:
        sim.advance()
    sim.close()
    files = os.listdir(str(tmp_path))
    cdm = sorted(f for f in files if f.startswith("pofk_") and f.endswith("_CDM.txt"))
    tot = sorted(f for f in files if f.startswith("pofk_") and f.endswith("_total.txt"))
    assert len(cdm) == len(tot) >= 3, files
    assert [f[:-len("CDM.txt")] for f in cdm] == [f[:-len("total.txt")] for f in tot]
    f2 = ((hp.Omega - hp.OmegaNu) / hp.Omega) ** 2
    for a, b in zip(cdm, tot):
        pc = np.loadtxt(os.path.join(str(tmp_path), a), ndmin=2)
        pt = np.loadtxt(os.path.join(str(tmp_path), b), ndmin=2)
        np.testing.assert_array_equal(pc[:, 0], pt[:, 0])
        r = pt[:, 1] / pc[:, 1]
        assert np.all(r > f2 * 0.99) and np.all(r < 1.01), (a, r)
        assert r[-1] < r[0], (a, r)
