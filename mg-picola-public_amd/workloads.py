"""Paramfile values of the benchmark and parity workloads.

Each workload is the reference's paramfile interface filled in for one
configuration: synthetic Gaussian initial conditions from the Eisenstein & Hu
no-wiggle P(k) (WhichSpectrum = 2), Omega_m 0.3, h 0.7, n_s 0.96, sigma_8 0.8,
Omega_b 0.049, 2LPT initial conditions at z = 19, COLA time stepping to z = 0.
"""

BASE = dict(
    modified_gravity_active=0, use_lcdm_growth_factors=0, input_pofk_is_for_lcdm=0, input_sigma8_is_for_lcdm=0,
    inverted_initial_condition=0, amplitude_fixed_initial_condition=0, include_screening=0,
    pofk_compute_every_step=0, pofk_compute_rsd_pofk=0, pofk_nbins=0, pofk_bintype=0, pofk_subtract_shotnoise=1,
    pofk_kmin=0.0, pofk_kmax=0.0,
    fofr0=1e-5, nfofr=1.0, Rsmooth=1.0, rcH0_DGP=1.0, range_symm=1.0, assb_symm=0.5, beta_symm=1.0,
    wBD=100.0, Omegah2=0.147, Omegar2=0.0, Omegav2=0.343,
    OutputDir="output", FileBase="run", OutputRedshiftFile="output_redshifts.dat", NumFilesWrittenInParallel=1,
    UseCOLA=1, Buffer=1.5, Nmesh=128, Nsample=64, Box=256.0, Init_Redshift=19.0, Seed=5001, SphereMode=0,
    WhichSpectrum=2, WhichTransfer=0, FileWithInputSpectrum="none", FileWithInputTransfer="none",
    Omega=0.3, OmegaBaryon=0.049, HubbleParam=0.7, Sigma8=0.8, PrimordialIndex=0.96,
    UnitVelocity_in_cm_per_s=1e5, UnitLength_in_cm=3.085678e24, UnitMass_in_g=1.989e43,
    InputSpectrum_UnitLength_in_cm=3.085678e24,
)

# name: (model, values, nsteps to z = 0)
WORKLOADS = {
    # Plain LCDM COLA, 64^3 particles on a 128^3 mesh, Box 256 Mpc/h, 10 steps (LCDM via an f(R)
    # build with modified gravity off, as the reference runs LCDM)
    "lcdm_64_128": ("FOFR_LCDM", dict(Nmesh=128, Nsample=64, Box=256.0), 10),
    # f(R) Hu-Sawicki fR0 = 1e-5, n = 1, scale-dependent growth, screening, P(k) every step with 512
    # linear bins: 512^3 particles, 1024^3 mesh, Box 1024 Mpc/h, 30 steps
    "fofr_sd_512_1024": ("FOFR", dict(Nmesh=1024, Nsample=512, Box=1024.0, modified_gravity_active=1,
                                      include_screening=1, pofk_compute_every_step=1, pofk_nbins=512), 30),
    # nDGP (rc H0 = 1) with density screening: the realizations of the batched ensemble,
    # Box = 2 Np Mpc/h, Nmesh = 2 Np, 20 steps (sizes 96, 128 and 192 particles per side)
    "dgp_96_192": ("DGP", dict(Nmesh=192, Nsample=96, Box=192.0, modified_gravity_active=1, include_screening=1,
                               Rsmooth=1.0, rcH0_DGP=1.0), 20),
    "dgp_128_256": ("DGP", dict(Nmesh=256, Nsample=128, Box=256.0, modified_gravity_active=1, include_screening=1,
                                Rsmooth=1.0, rcH0_DGP=1.0), 20),
    "dgp_192_384": ("DGP", dict(Nmesh=384, Nsample=192, Box=384.0, modified_gravity_active=1, include_screening=1,
                                Rsmooth=1.0, rcH0_DGP=1.0), 20),
    # the same run on the coarse mesh of the learned multiplier (Nmesh = Nsample = 512)
    "fofr_sd_512_512": ("FOFR", dict(Nmesh=512, Nsample=512, Box=1024.0, modified_gravity_active=1,
                                     include_screening=1, pofk_compute_every_step=1, pofk_nbins=256), 30),
}


def values_for(name, outdir, **over):
    model, vals, nsteps = WORKLOADS[name]
    v = dict(BASE)
    v.update(vals)
    v.update(OutputDir=outdir, FileBase=name, OutputRedshiftFile=f"{outdir}/output_redshifts.dat")
    v.update(over)
    return model, v, nsteps
