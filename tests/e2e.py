"""End-to-end COLA runs in the CPU oracle for the parity tests.

The oracle step functions restate the reference's kernels; this module strings
them together with the reference's schedule (src/main.c:386-611) for runs with
one output at z = 0.  Growth factors, time integrals and model scalars come from
the host cosmology (pinned on its own by tests/test_host_cosmology.py), so a
comparison against the GPU run isolates the mesh and particle path.
"""
import numpy as np

import mgpicola as M
import oracle as O


def wrapf(x, box):
    """periodic_wrap on float32 (src/auxPM.c:649-655)."""
    b = np.float32(box)
    x = x.astype(np.float32)
    for _ in range(4):
        x = np.where(x >= b, (x - b).astype(np.float32), x)
        x = np.where(x < 0, (x + b).astype(np.float32), x)
    x = np.where(x == b, np.float32(0), x)
    return x.astype(np.float32)


def oracle_run(hp, cosmo, nsteps, pofk=None):
    """COLA run z_init -> 0 in nsteps steps; returns (pos, vel, sumV), and with
    pofk = (nbins, bintype, subtract, kmin, kmax) also the binned P(k) of every
    force evaluation as a list of (P, k_mean, modes)."""
    N, Ns, box = hp.Nmesh, hp.Nsample, hp.Box
    model = {M.MODELS["FOFR"]: O.FOFR, M.MODELS["FOFR_LCDM"]: O.FOFR, M.MODELS["DGP"]: O.DGP,
             M.MODELS["BRANSDICKE"]: O.BRANSDICKE, M.MODELS["MBETA"]: O.MBETA}[hp.model]
    c = O.cfg(N, Ns, box, hp.Omega, model=model, mg_active=hp.modified_gravity_active,
              include_screening=hp.include_screening,
              filter=O.FILTER_GAUSSIAN if hp.model == M.MODELS["DGP"] else O.FILTER_NONE)
    A = 1.0 / (1.0 + hp.Init_Redshift)
    Di, Di2 = cosmo.growth(M.G_D, A), cosmo.growth(M.G_D2, A)
    cd = O.ic_delta(c, hp.Seed, cosmo.ic_pk_table(), hp.SphereMode, hp.amplitude_fixed_initial_condition,
                    hp.inverted_initial_condition)
    sd = bool(hp.scale_dependent)
    q = (np.indices((Ns, Ns, Ns)).reshape(3, -1).astype(np.float64)) * (box / Ns)
    if sd:
        (cd2,) = O.ic_displacements(c, cd, 1)

        def fields(field, a0, a1):
            g1, g2 = cosmo.lpt_tables(field, a0, a1)
            return O.lpt_field(c, cd, g1), O.lpt_field(c, cd2, g2)

        za, lp = fields(M.FIELD_D, A, A)
        pos = wrapf((q + za.astype(np.float64) + lp.astype(np.float64)).astype(np.float32), box)
    else:
        za, lp = O.ic_displacements(c, cd, 0)
        pos = wrapf((q + za.astype(np.float64) * Di + lp.astype(np.float64) * Di2).astype(np.float32), box)
    vel = np.zeros_like(pos)
    AI = AF = A
    da = (1.0 - A) / nsteps
    sV = np.zeros(3)
    pks = []
    for i in range(nsteps + 1):
        final = i == nsteps
        AF = A if final else A + 0.5 * da
        AFF = A + (0.0 if final else da)
        d = O.deposit(c, pos)
        dx = d.copy()
        O.fft_r2c_3d(N, d)
        if pofk is not None:
            nb, bt, sub, kmin, kmax = pofk
            pks.append(O.pofk(c, d, nb, bt, kmin, kmax, sub))
        phik = None
        if hp.modified_gravity_active:
            m = cosmo.mg(A)
            mg = O.MG(m.a, m.coupling, m.mass2, m.phi_crit, m.dgp_fac, m.rsmooth, m.geff)
            phik = O.fifth_force(c, mg, d, dx)
        f = O.forces(c, d, phik)
        disp, sD = O.gather(c, pos, f)
        dda = cosmo.Sphi(AI, AF, A)
        if sd:
            dd1, dd2 = fields(M.FIELD_deltaD, A, AFF)
            k1, k2 = fields(M.FIELD_ddDddy, A, AFF)
            sV = O.kick(vel, disp, k1, k2, 0, 0, 1, hp.Omega, dda, A, 1, sD)
        else:
            sV = O.kick(vel, disp, za, lp, cosmo.growth(M.G_ddDddy, A), cosmo.growth(M.G_ddD2ddy, A), 0, hp.Omega,
                        dda, A, 1, sD)
        if final:
            break
        dyyy = cosmo.Sq(A, AFF, AF)
        if sd:
            O.drift(pos, vel, dd1, dd2, 0, 0, 1, box, dyyy, 1, sV)
        else:
            O.drift(pos, vel, za, lp, cosmo.growth(M.G_D, AFF) - Di, cosmo.growth(M.G_D2, AFF) - Di2, 0, box, dyyy,
                    1, sV)
        A, AI = AFF, AF
        Di, Di2 = cosmo.growth(M.G_D, A), cosmo.growth(M.G_D2, A)
    if pofk is not None:
        return pos, vel, sV, pks
    return pos, vel, sV


def periodic_diff(a, b, box):
    d = np.asarray(a, np.float64) - np.asarray(b, np.float64)
    return d - box * np.round(d / box)
