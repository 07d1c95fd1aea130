/*
 * mgpu.h — C ABI of the MI355X (gfx950) implementation of MG-PICOLA's
 * particle-mesh COLA step and 2LPT initial-condition FFTs.
 *
 * Plain C: opaque context, plain pointers and sizes, int return codes.
 * One context owns every device buffer of one simulation (particles, meshes,
 * FFT tables) and one HIP stream; calls on a context are serialised on that
 * stream and must come from one host thread.  Host pointers passed in are
 * read/written synchronously before the call returns.
 *
 * Each entry point replaces one global-state routine of the reference
 * (cited below as file:line in HAWinther/MG-PICOLA-PUBLIC).  The reference
 * reports errors by FatalError -> MPI_Abort (src/auxPM.c:668-673); here every
 * call returns MGPU_OK (0) or a negative MGPU_E_* code and mgpu_last_error()
 * describes the failure, so a host wrapper can keep the abort behaviour.
 *
 * Layouts:
 *   particles  SoA float, x[0..np), y[0..np), z[0..np) concatenated ([3][np]);
 *              particle i is Lagrangian lattice point i = (ix*Ns + iy)*Ns + iz
 *              (src/main.c:257-309) and is never reordered.
 *   meshes     returned to the host in the reference's FFTW layout without the
 *              ghost slab: real [N][N][N], complex [N][N][N/2+1] (re,im).
 *   per-|n|^2 tables  t[n2], n2 = nx^2+ny^2+nz^2 over integer wavenumbers,
 *              length mgpu_n2_table_len(N) = 3*(N/2)^2 + 1.
 */
#ifndef MGPU_H
#define MGPU_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MGPU_OK 0
#define MGPU_E_INVALID (-1)     /* bad argument or unsupported size */
#define MGPU_E_NOMEM (-2)       /* device allocation failed */
#define MGPU_E_HIP (-3)         /* HIP runtime error */
#define MGPU_E_STATE (-4)       /* call out of order (e.g. forces before density) */
#define MGPU_E_NODEVICE (-5)    /* no GPU visible */

/* Gravity models: the compile-time MODEL= switch of the reference (Makefile:33-110). */
enum { MGPU_LCDM = 0, MGPU_FOFR = 1, MGPU_DGP = 2, MGPU_BRANSDICKE = 3, MGPU_MBETA = 4 };
/* Density-screening smoothing filter (src/mg.h:318-338). */
enum { MGPU_FILTER_NONE = 0, MGPU_FILTER_GAUSSIAN = 1, MGPU_FILTER_TOPHAT = 2, MGPU_FILTER_SHARPK = 3 };
/* Mesh ids for mgpu_get_mesh / mgpu_set_mesh. */
enum {
  MGPU_MESH_DENSITY_X = 0,  /* delta(x) after PtoMesh's deposit (real); the r2c (without MG) or the force
                               pass (with MG) reuse its memory, so read it before them */
  MGPU_MESH_DENSITY_K = 1,  /* delta(k) after the r2c (complex), P3D */
  MGPU_MESH_PHI_K = 2,      /* fifth-force potential phi(k) (complex), P3D_mgarray_two; until mgpu_forces */
  MGPU_MESH_FORCE_X = 3,    /* force meshes after the c2r (real), N11 N12 N13 */
  MGPU_MESH_FORCE_Y = 4,
  MGPU_MESH_FORCE_Z = 5,
  MGPU_MESH_CDELTA1 = 6,    /* stored 1LPT delta(k) (complex), cdelta_cdm */
  MGPU_MESH_CDELTA2 = 7     /* stored 2LPT -source(k) (complex), cdelta_cdm2 */
};
/* Per-particle LPT slots.  Not scale-dependent: slot 0 = P.D (ZA), slot 1 = P.D2
 * (2LPT).  Scale-dependent: slot 0 = D or ddDddy field, slot 1 = dDdy or deltaD
 * field, each the sum of the 1LPT and 2LPT fields (src/vars.h:305-312). */
enum { MGPU_LPT_SLOT0 = 0, MGPU_LPT_SLOT1 = 1 };

typedef struct {
  int nmesh;               /* Nmesh: 2^a 3^b, 8 <= Nmesh <= 2048 */
  int nsample;             /* Nsample: particles per dimension */
  double box;              /* Box [Mpc/h] */
  double omega;            /* Omega_m */
  int model;               /* MGPU_LCDM ... */
  int mg_active;           /* modified_gravity_active */
  int include_screening;
  int filter;              /* MGPU_FILTER_* for DGP smoothing */
  int scale_dependent;     /* SCALEDEPENDENT build */
  int use_cola;            /* UseCOLA */
  /* P(k) binning after adjust_pofk_parameters (src/compute_pofk.c:758-805);
   * kmin/kmax in integer-wavenumber units */
  int pofk_nbins, pofk_bintype, pofk_subtract_shotnoise;
  double pofk_kmin, pofk_kmax;
} mgpu_params;

/* Scalars of the model hooks at a = aexp_global, evaluated by the host
 * (src/user_defined_functions.h:413-773). */
typedef struct {
  double a;
  double coupling;         /* coupling_function(a) */
  double mass2;            /* mass2_of_a(a) = m^2(a)/H0^2 */
  double phi_crit;         /* screening threshold of screening_factor_potential at a */
  double dgp_fac;          /* 8/9 Omega (rcH0/beta_DGP(a))^2 (screening_factor_density) */
  double rsmooth;          /* Rsmooth [Mpc/h] */
  double geff;             /* GeffoverG(a, 0), JBD */
} mgpu_mg;

/* Binned power spectrum as compute_power_spectrum forms it (src/compute_pofk.c:71-236). */
typedef struct {
  double *pofk;            /* [nbins] normalised P(k) [(Mpc/h)^3], shot noise handled */
  double *kmean;           /* [nbins] mean |k| of the bin [h/Mpc] */
  double *nmodes;          /* [nbins] mode count (weights 1 or 2) */
} mgpu_pofk;

size_t mgpu_n2_table_len(int nmesh);
const char *mgpu_last_error(const void *ctx_or_null);
int mgpu_device_info(char *name, size_t name_len, size_t *global_mem, int *cu_count, size_t *lds_per_block);

/* Device memory of a context, by part: mgpu_create allocates the sum of all
 * parts but MGPU_MEM_RSD (total), and fails with MGPU_E_NOMEM, before
 * allocating, when that exceeds the free device memory; the RSD records are
 * allocated by the first mgpu_rsd_power_spectrum.  The reference allocates and
 * frees its meshes per step instead (MEMORY_MODE, src/auxPM.c:46-102).
 * parts: MGPU_MEM_PARTS sizes in bytes, or NULL.  Needs no device. */
enum {
  MGPU_MEM_PARTICLES = 0, /* Pos, Vel, Disp and the two LPT slots: 60 bytes per particle */
  MGPU_MEM_MESHES,        /* the mesh slots of the step */
  MGPU_MEM_DEPOSIT,       /* tile bins and particle records of the deposit and gather */
  MGPU_MEM_STORED,        /* scale-dependent: the stored delta(k) and 2LPT source (cdelta_cdm, cdelta_cdm2) */
  MGPU_MEM_LATTICE,       /* scale-dependent with Nmesh % Nsample == 0: the folded lattice spectra */
  MGPU_MEM_TABLES,        /* reduction scratch, P(k) bins and windows, per-|n|^2 tables, twiddles */
  MGPU_MEM_RSD,           /* redshift-space particle records (first RSD call) */
  MGPU_MEM_PARTS
};
int mgpu_memory_plan(const mgpu_params *p, size_t *parts, size_t *total);

/* ---- context; replaces initialize_ffts/initialize_parts + the MEMORY_MODE
 *      allocations of GetDisplacements (src/2LPT.c:47-176, src/auxPM.c:46-102) ---- */
typedef struct mgpu_ctx mgpu_ctx;
int mgpu_create(const mgpu_params *p, mgpu_ctx **out);
void mgpu_destroy(mgpu_ctx *ctx);
int mgpu_sync(mgpu_ctx *ctx);

/* ---- particle state (struct part_data, src/vars.h:293-353) ---- */
int mgpu_set_particles(mgpu_ctx *ctx, const float *pos, const float *vel);
int mgpu_get_particles(mgpu_ctx *ctx, float *pos, float *vel);
int mgpu_set_lpt(mgpu_ctx *ctx, int slot, const float *d);
int mgpu_get_lpt(mgpu_ctx *ctx, int slot, float *d);
int mgpu_get_disp(mgpu_ctx *ctx, float *disp);            /* Disp[3] of the last MtoParticles */

/* ---- PM step ---- */
/* PtoMesh: CIC deposit (deterministic 64-bit fixed point), delta(x) copy for
 * the MG solver, r2c, and P(k) of delta(k) when pk != NULL.  src/auxPM.c:280-432 */
int mgpu_pm_density(mgpu_ctx *ctx, mgpu_pofk *pk);
/* The two halves of mgpu_pm_density: the CIC deposit alone (delta(x) stays in
 * real layout, readable as MGPU_MESH_DENSITY_X; src/auxPM.c:292-362), and the
 * r2c + P(k) of the deposited field (src/auxPM.c:365-372). */
int mgpu_pm_deposit(mgpu_ctx *ctx);
int mgpu_pm_transform(mgpu_ctx *ctx, mgpu_pofk *pk);
/* compute_power_spectrum of the current delta(k).  src/compute_pofk.c:71 */
int mgpu_power_spectrum(mgpu_ctx *ctx, mgpu_pofk *pk);
/* ComputeFifthForce: f(R)/m-beta potential screening, DGP density screening,
 * JBD Geff(a).  src/user_defined_functions.h:687-719, src/mg.h:21-260 */
int mgpu_fifth_force(mgpu_ctx *ctx, const mgpu_mg *mg);
/* Forces: Green's function -i n/n^2 and 3 c2r.  src/auxPM.c:437-555 */
int mgpu_forces(mgpu_ctx *ctx);
/* MtoParticles: CIC gather into Disp, mean over particles.  src/auxPM.c:560-644 */
int mgpu_mesh_to_particles(mgpu_ctx *ctx, double sumD[3]);
/* RSD multipoles of one line of sight (axis 1 = y, 2 = z): PtoMesh_RSD, the r2c and
 * bin_up_RSD_power_spectrum of compute_RSD_powerspectrum (src/compute_pofk.c:279-395,
 * :403-455, :496-753).  Positions move along the line of sight by vnorm * V, vnorm =
 * (Hubble/A)/(100 A hubble(A)) * Nmesh/Box; V = Vel plus, with COLA, the LPT velocity:
 * scale-dependent, LPT slot 1 must hold dD/dy (both orders, FIELD_dDdy); otherwise
 * slot0 * s1 + slot1 * s2 with s1 = dDdy(A), s2 = dD2dy(A).  The binning is the P(k)'s
 * (pofk_* parameters of the context).  P0, P2, P4 (shot noise subtracted when
 * pofk_subtract_shotnoise), k = the bin's nominal k, n = modes; each array nbins long,
 * any may be NULL.  Overwrites the force meshes. */
typedef struct {
  double *P0, *P2, *P4, *k, *n;
} mgpu_rsd_pofk;
int mgpu_rsd_power_spectrum(mgpu_ctx *ctx, int axis, double vnorm, double s1, double s2, mgpu_rsd_pofk *out);
/* GetDisplacements = PtoMesh + ComputeFifthForce + Forces + MtoParticles.
 * mg may be NULL when modified gravity is off.  src/auxPM.c:37-103
 * For LCDM and the screened f(R), m-beta and DGP solvers the x passes run in
 * pairs (forward x, multiplier, inverse x in one kernel), so phi(k) is never
 * stored: MGPU_MESH_PHI_K is only defined after mgpu_fifth_force.  The other
 * meshes, Disp and the P(k) are those of the four separate calls. */
int mgpu_get_displacements(mgpu_ctx *ctx, const mgpu_mg *mg, mgpu_pofk *pk, double sumD[3]);
/* Kick: v += dda [-1.5 Omega (Disp - sumD) - UseCOLA lpt/A]; returns the new
 * mean velocity.  Not scale-dependent lpt = slot0*s1 + slot1*s2 with
 * s = (ddDddy, ddD2ddy); scale-dependent lpt = slot0.  src/main.c:688-742 */
int mgpu_kick(mgpu_ctx *ctx, double dda, double A, double s1, double s2, const double sumD[3], double sumV[3]);
/* Drift: x += (v - sumV) dyyy + UseCOLA lpt, periodic wrap.  Not
 * scale-dependent lpt = slot0*s1 + slot1*s2 with s = (deltaD, deltaD2);
 * scale-dependent lpt = slot1.  src/main.c:747-787 */
int mgpu_drift(mgpu_ctx *ctx, double dyyy, double s1, double s2, const double sumV[3]);
/* Kick then Drift in one pass, for the steps with no output between them
 * (src/main.c:688-787).  sumV: in, the mean velocity before the kick, which the
 * drift uses (the kick term has zero mean, so it is the mean after the kick up
 * to rounding); out, the mean after the kick.  Disp is left as gathered (not
 * reduced by sumD), so a further mgpu_kick needs a new mgpu_mesh_to_particles. */
int mgpu_kick_drift(mgpu_ctx *ctx, double dda, double A, double s1k, double s2k, const double sumD[3], double dyyy,
                    double s1d, double s2d, double sumV[3]);
/* One COLA step between outputs that never waits for the device: the
 * GetDisplacements of mgpu_get_displacements (without P(k)) and the Kick and
 * Drift of mgpu_kick_drift, with the means of Disp and of the last kick's
 * velocities kept in device memory (the reference's sumDxyz and sumxyz,
 * src/auxPM.c:635-644, src/main.c:735-740) instead of coming back to the host.
 * With graph != 0 the step's launches are captured in a hipGraph on each call
 * (after one eager call) and the context's executable graph is updated in place
 * with this step's scalars and launched: one launch per step.  Capture needs
 * the process to itself: while another host thread makes HIP calls (several
 * contexts stepping at once), the runtime invalidates the capture and the call
 * fails with MGPU_E_HIP; use graph = 0 there.  Results are bit-identical to
 * mgpu_get_displacements + mgpu_kick_drift.  For runs whose
 * step is one fused sequence: not scale-dependent, no P(k) that step, no
 * learned multiplier or neutrino term, timers off; the velocity sums must be on
 * the device (any mgpu_kick or mgpu_kick_drift since the particles were set).
 * Other cases return MGPU_E_STATE.  Read the means with mgpu_get_means. */
int mgpu_step(mgpu_ctx *ctx, const mgpu_mg *mg, double dda, double A, double s1k, double s2k, double dyyy,
              double s1d, double s2d, int graph);
/* The device-resident means: Disp of the last gather, velocities of the last kick. */
int mgpu_get_means(mgpu_ctx *ctx, double sumD[3], double sumV[3]);
/* A batch of count realizations of one run configuration (the paired,
 * fixed-amplitude ICs of an ensemble: seeds and inversion differ, the mesh, the
 * lattice, the box, the model and the time steps do not).  members[0..count-1]
 * are full contexts, each usable with every call of this header and destroyed
 * with mgpu_destroy; they share one HIP stream and one device allocation, in
 * which member r's buffers lie at a fixed stride from member 0's.  Each member
 * runs the main loop of src/main.c:394-611 as a solo run would. */
int mgpu_batch_create(const mgpu_params *p, int count, mgpu_ctx **members);
/* mgpu_ic_displacement_fields for every member of a batch at once (members in
 * the order of mgpu_batch_create; seeds[r] and inverted[r] per member, the
 * sphere mode, the amplitude fixing and pk_n2 shared): each kernel of the IC is
 * one launch over all members.  Each member's LPT slots are bit-identical to
 * mgpu_ic_displacement_fields on it alone.  Not scale-dependent runs
 * (MGPU_E_STATE).  src/2LPT.c:185-495 */
int mgpu_batch_ic_displacement_fields(mgpu_ctx *const *members, int count, const unsigned long *seeds,
                                      int sphere_mode, int amplitude_fixed, const int *inverted,
                                      const double *pk_n2);
/* mgpu_step for every member of a batch at once (members in the order of
 * mgpu_batch_create): each kernel of the step is one launch over all members
 * (grid z = member), so small meshes fill the device.  Every member must meet
 * mgpu_step's conditions; the result of each member is bit-identical to
 * mgpu_step on it alone. */
int mgpu_batch_step(mgpu_ctx *const *members, int count, const mgpu_mg *mg, double dda, double A, double s1k,
                    double s2k, double dyyy, double s1d, double s2d, int graph);

/* ---- initial conditions ---- */
/* displacement_fields (Gaussian): seed table + per-column RANLUX stream,
 * delta(k) at z = 0, ZA and 2LPT displacements.  Not scale-dependent: fills
 * LPT slots 0/1 with ZA and 2LPT lattice displacements (mean removed).
 * Scale-dependent: keeps delta(k) and -source(k) (cdelta_cdm, cdelta_cdm2).
 * pk_n2[n2] = PowerSpec(2 pi/Box sqrt(n2)) including any MG rescaling.
 * src/2LPT.c:185-495, src/2LPT.c:1204-1520 */
int mgpu_ic_displacement_fields(mgpu_ctx *ctx, unsigned long seed, int sphere_mode, int amplitude_fixed,
                                int inverted, const double *pk_n2);
/* mgpu_ic_displacement_fields without any device allocation.  No counterpart in
 * the reference: it is for callers that cannot allocate device memory (code
 * under HIP graph capture, serving loops, harnesses that forbid hipMalloc).
 * The per-call seed table (Nmesh^2 unsigned) is staged in a 16 MiB scratch array
 * that each GPU holds until the process ends (at most Nmesh = 2048; larger
 * tables give MGPU_E_INVALID).  Every call on one GPU shares that GPU's
 * scratch: calls that overlap on one GPU (contexts on several host threads)
 * race and must be serialised by the caller.  Results are identical to
 * mgpu_ic_displacement_fields. */
int mgpu_ic_displacement_fields_noalloc(mgpu_ctx *ctx, unsigned long seed, int sphere_mode, int amplitude_fixed,
                                        int inverted, const double *pk_n2);
/* Particle init: Vel = 0 (COLA) and Pos = q + D*Di + D2*Di2 (scale-dependent:
 * q + slot0), periodic wrap.  src/main.c:226-309 */
int mgpu_init_particles(mgpu_ctx *ctx, double Di, double Di2);
/* from_cdisp_store_to_ZA + assign_displacment_field_to_particles for both LPT
 * orders at once: slot = field(cdelta1, g1) + field(cdelta2, g2), each field
 * i k/k^2 delta(k) g(|k|) c2r'ed, read at the lattice and mean-removed.
 * g1_n2/g2_n2 already include normfactor (1 and -3/7/N^3).
 * src/2LPT.c:1539-1735, src/2LPT.c:1758-2005 */
int mgpu_lpt_fields(mgpu_ctx *ctx, int slot, const double *g1_n2, const double *g2_n2);

/* Growth factor of one LPT order as a function of ln k, as the restriction of
 * the reference's bicubic (ln a, ln k) spline to the current a: cubic Hermite
 * pieces on nk uniform ln k nodes, g(k) = norm * num(ln k) / den(ln k), ln k
 * clamped to the node range (Lookup_GSL_2D_Spline, src/wrappers.c:129-137;
 * growth_X_scaledependent, src/cosmo.c:987-1010). */
typedef struct {
  double lnk0, dlnk;
  int nk;
  const double *num, *dnum;  /* numerator value and d/dln k at the nodes */
  const double *den, *dden;  /* denominator (growth at a = 1) value and derivative */
  double norm;               /* normfactor: 1 (1LPT) or -3/7/N^3 (2LPT) */
} mgpu_kslice;
/* mgpu_lpt_fields with the growth evaluated on the GPU from the two slices. */
int mgpu_lpt_fields_slices(mgpu_ctx *ctx, int slot, const mgpu_kslice *o1, const mgpu_kslice *o2);
/* Both LPT slots in one call: the four assign_displacment_field_to_particles
 * calls that precede Kick (src/main.c:496-501: slot 0 from s0o1 + s0o2, slot 1
 * from s1o1 + s1o2).  The stored spectra are read once for both slots. */
int mgpu_lpt_fields_slices2(mgpu_ctx *ctx, const mgpu_kslice *s0o1, const mgpu_kslice *s0o2,
                            const mgpu_kslice *s1o1, const mgpu_kslice *s1o2);
/* Massive neutrinos (FOFRNU): PtoMesh's neutrino term, src/auxPM.c:383-420,
 *   delta(k) <- cdmfac delta(k) + fac(k) cdelta_IC(k)   for every mode but k = 0,
 * cdmfac = (Omega - OmegaNu)/Omega, fac = OmegaNu/Omega N^3 T_nu(k, a)/T_cb(k, 1) given as a
 * ln k slice (norm num/den, as mgpu_kslice).  Set once per step before the density;
 * mgpu_pm_density and mgpu_get_displacements then apply it after the CDM P(k).  fac ==
 * NULL turns it off.  Needs a scale-dependent context (the IC spectrum is kept). */
int mgpu_set_neutrinos(mgpu_ctx *ctx, double cdmfac, const mgpu_kslice *fac);
/* the neutrino term alone, on the current delta(k) */
int mgpu_neutrino_response(mgpu_ctx *ctx);


/* ---- inspection (tests, output) ---- */
int mgpu_get_mesh(mgpu_ctx *ctx, int mesh, double *host);
int mgpu_set_mesh(mgpu_ctx *ctx, int mesh, const double *host);

/* ---- per-phase device time of the last calls, ms (timer categories of src/timer.h:11-29) ---- */
typedef struct {
  double ptomesh, fft, pofk, modified_gravity, forces, mtoparticles, kick, drift, lpt_fields, ic;
} mgpu_timers;
int mgpu_get_timers(mgpu_ctx *ctx, mgpu_timers *t, int reset);
int mgpu_enable_timers(mgpu_ctx *ctx, int on);

/* ---- scale-dependent growth tables; replaces calculate_scale_dependent_growth_factor
 *      (src/cosmo.c:798-981): the ODEs ode_growth_D / ode_growth_DLCDM (src/cosmo.c:206-261)
 *      with the hooks GeffoverG, second_order_kernel(k, k/sqrt2, k/sqrt2, 0, a) and
 *      GeffoverG_neutrino_1LPT/2LPT (src/user_defined_functions.h:618-680, :800-904),
 *      one k per GPU thread, FP64, Dormand-Prince 5(4) at tolerance 1e-11 ---- */
typedef struct {
  int nx, ny;                       /* knots in x and in y */
  const double *x, *y;              /* increasing knots */
  const double *z, *zx, *zy, *zxy;  /* value, d/dx, d/dy, d2/dxdy at (x_i, y_j), index j*nx + i */
} mgpu_spline2d;

typedef struct {
  int model;      /* MGPU_LCDM, MGPU_FOFR or MGPU_MBETA (the scale-dependent models) */
  int mg_active;  /* modified_gravity_active and not use_lcdm_growth_factors: GeffoverG and gamma2 on */
  int lcdm;       /* 1: ode_growth_DLCDM (mu = 1, no gamma2) and its source terms */
  double omega;
  double fofr0, nfofr;                      /* f(R) Hu-Sawicki */
  double assb_symm, beta_symm, range_symm;  /* m(a), beta(a) of the symmetron-like model */
  double omega_nu, omega_b, omega_cdm;      /* omega_nu > 1e-6 turns the neutrino factors on */
  mgpu_spline2d t_nu, t_cdm, t_b;           /* CAMB transfer functions in (ln k, z) */
} mgpu_growth_model;

/* Integrates from x0 = ln a_ini for every ln k of lnk[0..nk) and samples the
 * solution at x0 + i dx, i < na.  out: 6 nk na doubles, index (f nk + j) na + i,
 * fields f: D1, dD1/dy (= dD1/dx Q/a), 1.5 Om a mu mu_nu1 D1, D2, dD2/dy,
 * 1.5 Om a mu mu_nu2 (D2 - D1^2) (mu = mu_nu = 1 for lcdm).  Synchronous; uses
 * the current device.  MGPU_E_INVALID if an ODE fails (NaN or 10^6 steps). */
int mgpu_growth_tables(const mgpu_growth_model *m, double x0, double dx, int na, const double *lnk, int nk,
                       double *out);

/* ---- learned Green's-function multiplier (no counterpart in the reference: it
 *      fills the grid_corr slot that Forces leaves at 1, src/auxPM.c:488-498) ----
 * A 4 -> 64 -> 64 -> 64 -> 1 GELU network maps x = (|k|/k_Nyquist, a, log10 f_R0,
 * s) to ln M; s is the rms over the mesh of ln(1 + delta) (floored at ln 0.01)
 * after the deposit; the table takes |k|/k_Nyquist clamped at 1.  Parameters (MGPU_MLP_NPARAMS floats, FP32 master copy):
 * W1[64][4], b1[64], W2[64][64], b2[64], W3[64][64], b3[64], W4[64], b4; W[out][in].
 * The forward pass rounds weights to BF16 and runs the 64x64 layers on MFMA with
 * BF16 activations and FP32 accumulation.  Sample arrays are host memory: x n*4,
 * targets t and weights w n floats; the loss is sum w (y - t)^2 / sum w. */
#define MGPU_MLP_NPARAMS 8705
typedef struct mgpu_mlp mgpu_mlp;
int mgpu_mlp_create(unsigned long long seed, mgpu_mlp **out); /* W ~ N(0, 1/fan_in), output layer 0: M = 1 */
void mgpu_mlp_destroy(mgpu_mlp *m);
int mgpu_mlp_get_params(mgpu_mlp *m, float *params);
int mgpu_mlp_set_params(mgpu_mlp *m, const float *params); /* also resets the Adam state */
int mgpu_mlp_forward(mgpu_mlp *m, const float *x, int n, float *y);
int mgpu_mlp_gradient(mgpu_mlp *m, const float *x, const float *t, const float *w, int n, float *grad, double *loss);
/* iters full-batch Adam steps (beta 0.9, 0.999, eps 1e-8); loss of the last evaluation */
int mgpu_mlp_train(mgpu_mlp *m, const float *x, const float *t, const float *w, int n, int iters, float lr,
                   double *loss);
/* From the next deposit on: compute s after every deposit and, with m != NULL,
 * multiply the force Green's function by M(|n|^2; a, theta, s).  a < 0 turns
 * both off.  The context keeps m (not owned) until it is replaced or turned off. */
int mgpu_set_learned_multiplier(mgpu_ctx *ctx, mgpu_mlp *m, double a, double theta);
int mgpu_density_feature(mgpu_ctx *ctx, double *s); /* s of the last deposit */

/* ---- FFT schedule (no counterpart in the reference: FFTW plans its own) ----
 * The strided column passes work on tiles of 8 or 16 columns; the width changes
 * no column's arithmetic (results are identical), only speed.  which: 0 the
 * Nmesh transforms (and every pass fused with them), 1 the Nsample lattice
 * transforms of the scale-dependent LPT fields. */
int mgpu_set_fft_tile(mgpu_ctx *ctx, int which, int tcols); /* 16 only where the tile fits a workgroup */
int mgpu_get_fft_tile(mgpu_ctx *ctx, int which, int *tcols);
/* Measure both widths on this device (forward + inverse y pass, best of three)
 * and keep the faster, for both plans.  It uses the force meshes and the
 * lattice spectra as scratch: call it before the initial conditions.
 * ms: 4 doubles (main 8, main 16, lattice 8, lattice 16; -1 where not run). */
int mgpu_autotune_fft(mgpu_ctx *ctx, double ms[4]);

/* ---- per-kernel HIP-event timing (no counterpart in the reference) ----
 * When enabled, every launch of the listed kernel classes is bracketed by HIP
 * events on the context's stream; mgpu_kernel_timing synchronises and reports
 * launches, summed device time and summed algorithmic bytes (compulsory HBM
 * reads + writes) since the last reset. */
/* Classes (each one kernel or one kernel family): deposit tile accumulation,
 * the mean-displacement sum in particle order, z-pass r2c / c2r, column (y and x) passes forward /
 * inverse of the Nmesh transforms (the Nsample-lattice transforms count as
 * MGPU_K_LPT), P(k), ..., the FFT passes with fused multipliers (Green's force
 * x pass, dual y pass, z screening pass, x pass with the MG epilogue), the
 * deposit's binning and sort, the lattice fold, the potential/smoothing x pass,
 * the density x pass with P(k) binning, and the two x-pass pairs of the fused
 * step (density forward x + potential/smoothing or force inverse x; effective
 * density forward x + MG epilogue + force inverse x). */
enum {
  MGPU_K_DEPOSIT = 0, MGPU_K_MEAN, MGPU_K_FFT_R2C, MGPU_K_FFT_C2R, MGPU_K_FFT_COL_FWD, MGPU_K_FFT_COL_INV,
  MGPU_K_POFK, MGPU_K_SPECTRAL, MGPU_K_REAL, MGPU_K_GATHER, MGPU_K_KICK, MGPU_K_DRIFT, MGPU_K_LPT, MGPU_K_READOUT,
  MGPU_K_FFT_GREEN, MGPU_K_FFT_DUAL, MGPU_K_FFT_SCREEN, MGPU_K_FFT_EPI, MGPU_K_DEP_SORT, MGPU_K_FOLD,
  MGPU_K_FFT_POT, MGPU_K_FFT_PK, MGPU_K_FFT_XDENS, MGPU_K_FFT_XFORCE, MGPU_K_COUNT
};
const char *mgpu_kernel_name(int kernel);
int mgpu_enable_kernel_timing(mgpu_ctx *ctx, int on);
int mgpu_kernel_timing(mgpu_ctx *ctx, int kernel, int *launches, double *ms, double *bytes, int reset);

#ifdef __cplusplus
}
#endif
#endif
