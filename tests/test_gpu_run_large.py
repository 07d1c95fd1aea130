"""The large single run of the learned multiplier's workload on one GPU:
f(R) f_R0 = 1e-5 with massive neutrinos (linear response, M_nu = 0.2 eV, the
reference's CAMB tables), 1024^3 particles on the learned-coarse 1024^3 mesh
in a 2048 Mpc/h box, 30 COLA steps from z = 19 to 0, CDM and total P(k) every
step and the RSD multipoles P0/P2/P4 at z = 0 (pofk_compute_rsd_pofk = 2), no
particle files; once plain and once with the learned Green's-function
multiplier (a network trained first on twelve 128^3 coarse/3x-fine pairs with
the same particle spacing and steps, as bench.py trains it).

No reference output exists at this size, so the runs are checked by what the
physics fixes: the memory plan fits the device; the P(k) files are finite and
positive; total/CDM lies between (1 - f_nu)^2 and 1 and falls with k; the
plain run's CDM P(k) at k < 0.02 h/Mpc (no shot-noise subtraction: the
initial lattice has none) grows from the first step (z = 19) to z = 0 by the
linear growth of an Omega_m = 0.3 LCDM within 12 % (the massive neutrinos
slow the CDM growth by a few per cent; f(R) acts below these scales); its redshift-space monopole over the real-space power and quadrupole
over monopole at k < 0.03 h/Mpc (sums over bins weighted by k^2, as their
mode counts) follow Kaiser's linear values for f = 0.3^0.55 within 12 % and
0.1.  The learned run's P(k) at z = 0 over the plain run's stays within 3 %
of 1 at k < 0.02 h/Mpc (the network is anchored to ln M = 0 at k = 0) and,
for every k < 1 h/Mpc, between 0.95 and 1/(1 - 0.30): the correction undoes
the coarse mesh's power deficit, which is 26 % at k = 1 on the 512^3 / 1024
Mpc/h benchmark against its 3x mesh, and removes none; the multiplier was
evaluated at every force evaluation.  The network is trained at 256 Mpc/h,
so this 2048 Mpc/h run also shows how far it carries in box size; its
accuracy against a 3x mesh is what bench.py measures."""
import math
import os
import time

import numpy as np
import pytest

import learned as LR
import mgpicola as M
import test_host as H
import workloads as W

pytestmark = pytest.mark.gpu

NP, BOX, NSTEPS = 1024, 2048.0, 30


def train_pair(tmp_path, seed, fofr0):
    """(coarse, 3x fine) training pair: 128^3 particles in 256 Mpc/h (the big run's particle spacing), the
    f(R) scale-dependent workload's 30 steps, LCDM when fofr0 is 0."""
    over = dict(Nsample=128, Box=256.0, Seed=seed, pofk_compute_every_step=1, pofk_nbins=64, pofk_kmin=0.0,
                pofk_kmax=math.pi * 128 / 256.0)
    if fofr0:
        over.update(fofr0=fofr0)
    else:
        over.update(modified_gravity_active=0, include_screening=0)
    out = []
    for nm in (128, 384):
        d = tmp_path / f"pair{seed}_{fofr0}_{nm}"
        d.mkdir()
        model, vals, _ = W.values_for("fofr_sd_512_1024", str(d), Nmesh=nm, **over)
        with open(vals["OutputRedshiftFile"], "w") as f:
            f.write("0, 30\n")
        pf = str(d / "params.txt")
        M.write_paramfile(pf, vals, model=model)
        hp = M.read_params(pf, model=model)
        hp.write_particles = 0
        out.append(hp)
    return tuple(out)


def linear_growth(a, om=0.3):
    """LCDM D(a) by the integral formula, normalised to a at early times."""
    from scipy import integrate
    ol = 1.0 - om

    def e(x):
        return math.sqrt(om / x ** 3 + ol)

    return 2.5 * om * e(a) * integrate.quad(lambda x: 1.0 / (x * e(x)) ** 3, 1e-8, a, epsabs=0, epsrel=1e-10)[0]


def run(tmp_path, tag, mlp):
    out = tmp_path / tag
    out.mkdir()
    outs = out / "outs.dat"
    outs.write_text(f"0, {NSTEPS}\n")
    hp, c = H.nu_cosmo(out, Nmesh=NP, Nsample=NP, Box=BOX, pofk_nbins=256, pofk_compute_every_step=1,
                       pofk_compute_rsd_pofk=2, pofk_subtract_shotnoise=0, OutputRedshiftFile=str(outs))
    c.close()
    hp.growth_on_gpu = 1
    hp.write_particles = 0
    t0 = time.time()
    sim = M.Sim(hp)
    sim.set_learned(mlp, True)
    sim.ic()
    sim.gpu.sync()
    print(f"{tag}: tables and initial conditions {time.time() - t0:.1f} s", flush=True)
    ts = []
    while not sim.finished:
        t = time.perf_counter()
        sim.advance()
        sim.gpu.sync()
        ts.append(time.perf_counter() - t)
        if len(ts) % 10 == 0:
            print(f"{tag}: step {len(ts)}, a = {sim.a:.4f}, {ts[-1] * 1e3:.0f} ms", flush=True)
    log = sim.log()
    sim.close()
    print(f"{tag}: {len(ts) - 1} steps, median {np.median(ts[1:-1]) * 1e3:.0f} ms/step", flush=True)
    files = os.listdir(str(out))
    cdm = sorted(f for f in files if f.startswith("pofk_") and f.endswith("_CDM.txt"))
    tot = sorted(f for f in files if f.startswith("pofk_") and f.endswith("_total.txt"))
    rsd = [f for f in files if f.startswith("pofk_RSD_")]
    assert len(cdm) >= NSTEPS - 1 and len(cdm) == len(tot) and len(rsd) == 1, (len(cdm), len(tot), rsd)
    z0 = [f for f in cdm if "_z0.000_" in f]
    assert z0, cdm[-3:]
    first = max(cdm, key=lambda f: float(f.split("_z")[-1].split("_")[0]))
    ld = lambda f: np.loadtxt(os.path.join(str(out), f), ndmin=2)
    return dict(hp=hp, log=log, pc=ld(z0[0]), pt=ld(z0[0].replace("_CDM.txt", "_total.txt")), p1=ld(first),
                zf=float(first.split("_z")[-1].split("_")[0]), rsd=ld(rsd[0]))


def low_k_sum(p, kmax):
    """k^2-weighted sum of P over the bins with 0 < k_mean < kmax (modes per bin grow as k^2)."""
    sel = (p[:, 2] > 0) & (p[:, 2] < kmax)
    return (p[sel, 2] ** 2 * p[sel, 1]).sum()


@pytest.mark.timeout(900)
def test_learned_coarse_1024_fofr_nu_rsd_run(tmp_path):
    parts, total = M.memory_plan(NP, NP, scale_dependent=1, pofk_nbins=256)
    gmem = M.device_info()["global_mem"]
    print(f"memory plan {total / 1e9:.1f} GB + RSD {parts['rsd'] / 1e9:.1f} GB of {gmem / 1e9:.1f} GB", flush=True)
    assert total + parts["rsd"] < 0.6 * gmem

    plain = run(tmp_path, "plain", None)
    mlp = M.MLP(seed=7)
    t0 = time.time()
    LR.train([train_pair(tmp_path, s, f) for s in (101, 102, 103) for f in (0.0, 1e-6, 1e-5, 1e-4)], mlp,
             iterations=30, iters=1500, lr=3e-3, eta=0.6, eta_decay=0.93, kmax_select=1.0, wpow=0.25, anchor=4.0,
             val_pairs=[train_pair(tmp_path, 104, f) for f in (0.0, 1e-5)])
    print(f"network trained in {time.time() - t0:.0f} s", flush=True)
    lrn = run(tmp_path, "learned", mlp)
    mlp.close()
    assert len(lrn["log"]) >= NSTEPS - 1 and all(np.isfinite(e[1]) for e in lrn["log"])

    hp = plain["hp"]
    f2 = ((hp.Omega - hp.OmegaNu) / hp.Omega) ** 2
    for r in (plain, lrn):
        pc, pt = r["pc"], r["pt"]
        assert np.all(np.isfinite(pc)) and np.all(pc[:, 1] > 0)
        ratio = pt[:, 1] / pc[:, 1]
        assert np.all(ratio > f2 * 0.99) and np.all(ratio < 1.01) and ratio[-1] < ratio[0], ratio
        assert np.all(np.isfinite(r["rsd"])) and np.all(r["rsd"][:, 1] > 0)

    g = low_k_sum(plain["pc"], 0.02) / low_k_sum(plain["p1"], 0.02)
    d2 = (linear_growth(1.0) / linear_growth(1.0 / (1.0 + plain["zf"]))) ** 2
    print(f"plain: P(k < 0.02) growth from z = {plain['zf']:.3f}: {g:.2f} against linear LCDM {d2:.2f}", flush=True)
    assert abs(g / d2 - 1.0) < 0.12

    rs, pc = plain["rsd"], plain["pc"]
    fg = 0.3 ** 0.55
    k0 = 1 + 2 * fg / 3 + fg ** 2 / 5
    q0 = (4 * fg / 3 + 4 * fg ** 2 / 7) / k0
    sel = rs[:, 0] < 0.03
    w = rs[sel, 0] ** 2
    mono = (w * rs[sel, 1]).sum() / (w * np.interp(rs[sel, 0], pc[:, 2], pc[:, 1])).sum()
    quad = (w * rs[sel, 2]).sum() / (w * rs[sel, 1]).sum()
    print(f"plain: Kaiser at k < 0.03: P0/P {mono:.3f} (linear {k0:.3f}), P2/P0 {quad:.3f} (linear {q0:.3f})",
          flush=True)
    assert abs(mono / k0 - 1.0) < 0.12
    assert abs(quad - q0) < 0.1

    sp = [e[1] for e in plain["log"]]
    sl = [e[1] for e in lrn["log"]]
    print(f"density feature s: first {sp[0]:.3f}, last {sp[-1]:.3f} (plain), {sl[-1]:.3f} (learned)", flush=True)
    ratio = lrn["pc"][:, 1] / plain["pc"][:, 1]
    kk = plain["pc"][:, 2]
    sel = (kk > 0) & (kk < 1.0)
    low = (kk > 0) & (kk < 0.02)
    at = [int(np.argmin(np.abs(kk - x))) for x in (0.01, 0.1, 0.25, 0.5, 0.75, 0.95)]
    print("learned / plain at z = 0: " + ", ".join(f"k {kk[i]:.3f}: {ratio[i]:.3f}" for i in at) +
          f"; max {ratio[sel].max():.3f}, min {ratio[sel].min():.3f}", flush=True)
    assert np.all(np.abs(ratio[low] - 1.0) < 0.03), ratio[low]
    assert np.all(ratio[sel] > 0.95) and np.all(ratio[sel] < 1.0 / (1.0 - 0.30)), ratio[sel]
