// Internal definitions shared by the translation units of libmgpu.so.
//
// Device mesh layout (one "slot" = N*N*HP complex64):
//   complex half spectrum  c[(ix*N + iy)*HP + kz],  kz in [0, N/2], HP = N/2+1
//                          rounded up to a multiple of 16 (128-byte rows);
//                          entries kz in (N/2, HP) are zero padding.
//   real mesh (in place)   r[(ix*N + iy)*2*HP + iz], iz in [0, N); the rest of
//                          each row is padding.
// This is FFTW's in-place r2c layout with a wider row so that every row starts
// on a 128-byte boundary and column tiles of 16 complex values are aligned.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string>
#include <type_traits>
#include <vector>

#include "../../include/mgpu.h"

#define MGPU_PI 3.14159265358979323846
#define MGPU_INVERSE_H0_MPCH 2997.92458
// mesh slots of a context
#define MGPU_NSLOTS 4
// seed-table entries of the allocation-free IC scratch: the Nmesh^2 columns of Nmesh = 2048
#define MGPU_IC_SCRATCH_SEEDS (2048 * 2048)
// columns per tile of the strided FFT passes (16: 128-byte row segments, 8: 64-byte)
#ifndef MGPU_FFT_TCOLS
#define MGPU_FFT_TCOLS 16
#endif

struct FFTPlan {
  int n = 0;       // transform size per dimension
  int hp = 0;      // padded complex row length
  int tcols = 16;  // columns per tile in the strided passes
  int kcat = -1;   // kernel-timing category of all passes (-1: by pass kind)
  float2 *tw = nullptr;  // W_N^m = exp(-2 pi i m / N), m < N
};

// One device allocation shared by the members of a batch (mgpu_batch_create):
// member r's per-realization buffers lie at base + r * stride, each at the same
// offset inside its share, so a kernel reaches realization blockIdx.z of a
// batched launch by adding blockIdx.z * stride to realization 0's pointers.
struct mgpu_arena {
  char *base = nullptr;
  size_t stride = 0;
  int count = 0;
  int refs = 0;                   // members not yet destroyed
  hipStream_t stream = nullptr;   // the one stream of all members
};

// realization blockIdx.z of a batched launch (zs = 0: an unbatched context).  Byte
// arithmetic on the pointer itself (not through an integer), so the compiler still
// sees a global-memory pointer and keeps global_load/global_store instead of flat
// accesses; a null pointer (an absent optional buffer) stays null.
template <class T>
__device__ __forceinline__ T *zbuf(T *p, size_t zs) {
  using B = typename std::conditional<std::is_const<T>::value, const char, char>::type;
  return reinterpret_cast<T *>(reinterpret_cast<B *>(p) + (p ? (size_t)blockIdx.z * zs : (size_t)0));
}

struct mgpu_ctx {
  mgpu_params p;
  int N = 0, Np = 0, HP = 0;
  size_t np = 0;          // particles
  size_t slot_elems = 0;  // complex elements per mesh slot
  hipStream_t stream = nullptr;
  int device = 0;
  size_t lds_max = 65536;

  // particles, SoA [3][np]
  float *pos = nullptr, *vel = nullptr, *disp = nullptr;
  float *lpt[2] = {nullptr, nullptr};

  // mesh slots (MGPU_NSLOTS, each N*N*HP complex64), reused through the step:
  //   0  delta(x) -> delta(k) (kept until the next deposit)
  //   1  delta(x) copy for the MG screening; then the force pair's G -> F_z
  //   2  MG potential work, phi(k); then i n_y G -> F_y
  //   3  i n_x G -> F_x
  // so the force meshes are slots {3, 2, 1} and delta(x) of an MG step and phi(k)
  // last until the force pass.  The initial conditions use slots 1-3 as work.
  float2 *slot[MGPU_NSLOTS] = {nullptr, nullptr, nullptr, nullptr};
  float2 *cdelta[2] = {nullptr, nullptr};
  bool have_phik = false;
  int dx_slot = 0;                  // slot holding delta(x) (1 when the MG screening needs it kept)
  int phi_slot = 2;                 // slot holding phi(k) after the fifth force
  int force_slot[3] = {3, 2, 1};    // slots of the x, y, z force meshes

  // reductions
  double *partial = nullptr;  // per-block partial sums
  size_t partial_len = 0;
  double *dsum = nullptr;     // final sums
  double *pk = nullptr;       // 3 * nbins
  double *winv = nullptr;     // P(k) deconvolution per axis: 1/sinc(pi d/N)^4 N^-2, d = 0..N/2
  float *winvf = nullptr;     // FP32 copy of winv
  int16_t *pk_bin = nullptr;  // P(k) bin of each |n|^2 (-1: outside the bins); null when nbins > 32767
  float4 *rsd_crd = nullptr;  // redshift-space particle records (allocated on the first RSD call)
  // massive neutrinos: PtoMesh's term delta = cdmfac delta + fac(|n|^2) cdelta_IC while nu_on
  bool nu_on = false;
  double nu_cdmfac = 1.0;
  double *nu_slice = nullptr;  // the factor's ln k slice (4 ny doubles)
  size_t nu_slice_len = 0;
  double *nu_fac = nullptr;    // factor per |n|^2 (n2len doubles)
  // learned Green's-function multiplier: while learn_on every deposit is followed by
  // the density feature (feat) and, with a network set, the table M(|n|^2) that the
  // force pass applies
  bool learn_on = false;
  mgpu_mlp *mlp = nullptr;
  double learn_a = 0.0, learn_theta = 0.0;
  double *feat_part = nullptr, *feat = nullptr;
  float *green_mult = nullptr;
  std::vector<double> pk_geom;  // mode-weighted |n| sums and mode counts per bin (2 * nbins), fixed per context
  double *n2tab[2] = {nullptr, nullptr};  // per-|n|^2 device tables
  float *n2tabf[2] = {nullptr, nullptr};  // FP32 copies of n2tab (lattice fold)
  size_t n2len = 0;

  FFTPlan fft;
  FFTPlan fft_lattice;  // Nsample-sized plan for folded LPT fields (when N % Ns == 0)
  float2 *lattice_slot = nullptr;  // 3 Nsample^3 half spectra
  double *slice_buf = nullptr;     // growth slices uploaded by mgpu_lpt_fields_slices
  size_t slice_len = 0;
  double *slice_host = nullptr;    // pinned staging buffer of the slices
  size_t slice_host_len = 0;
  hipEvent_t slice_ev = nullptr;   // the last upload from slice_host
  void *dep_scratch = nullptr;     // tiled deposit: bin keys, ranks, sorted indices, bin offsets
  size_t dep_scratch_bytes = 0;
  float4 *dep_rec = nullptr;       // particle records (x, y, z, index bits) in bin order (inside dep_scratch)
  unsigned *dep_off = nullptr;     // bin offsets of dep_rec, 8 bins (face classes) per tile
  bool sorted_valid = false;       // dep_rec holds the current positions
  unsigned dep_tiles_grid = 0;     // workgroups of the persistent tile deposit (0: not yet measured)
  // device-resident step (mgpu_step): dsum[0..2] the Disp sums of the last gather, dsum[3..5] the velocity
  // sums of the last kick; vsum_valid when dsum[3..5] matches the velocities
  bool vsum_valid = false;
  bool step_warm = false;           // one eager step has run (kernel attributes and launch shapes settled)
  hipGraphExec_t step_exec = nullptr;  // the per-step graph, updated in place each call
  unsigned *ic_scratch = nullptr;  // this device's seed-table scratch (allocation-free IC)
  // batches: the shared allocation (null for a context of its own), the bump pointer
  // while the context is created, and the launch shape of the step kernels: bz
  // realizations at byte stride zs (bz > 1 only inside mgpu_batch_step)
  mgpu_arena *arena = nullptr;
  char *arena_cur = nullptr, *arena_end = nullptr;
  size_t *arena_measure = nullptr;  // creation that only measures: the batch share's bytes are added here
  unsigned bz = 1;
  size_t zs = 0;

  // state
  bool have_deposit = false, have_density = false, have_forces = false, have_disp = false;

  // timers
  bool timers_on = false;
  mgpu_timers t{};
  hipEvent_t ev0 = nullptr, ev1 = nullptr;

  // per-kernel event timing
  bool ktimers_on = false;
  struct KRec {
    int cat;
    hipEvent_t a, b;
    double bytes;
  };
  std::vector<KRec> krec;
  int kcount[MGPU_K_COUNT] = {};
  double kms[MGPU_K_COUNT] = {}, kbytes[MGPU_K_COUNT] = {};

  std::string err;
};

// brackets one kernel launch with events when per-kernel timing is on
struct KScope {
  mgpu_ctx *ctx;
  mgpu_ctx::KRec r{};
  bool on = false;
  KScope(mgpu_ctx *c, int cat, double bytes);
  ~KScope();
};

// ---------------------------------------------------------------- helpers
int mgpu_fail(mgpu_ctx *ctx, int code, const std::string &msg);
int growth_tables(const mgpu_growth_model *m, double x0, double dx, int na, const double *lnk, int nk,
                  double *out);
// learned Green's-function multiplier (mlp.hip)
int mlp_create(unsigned long long seed, mgpu_mlp **out);
void mlp_destroy(mgpu_mlp *m);
int mlp_params(mgpu_mlp *m, float *get, const float *set);
int mlp_forward(mgpu_mlp *m, const float *x, int n, float *y);
int mlp_gradient(mgpu_mlp *m, const float *x, const float *t, const float *w, int n, float *grad, double *loss);
int mlp_train(mgpu_mlp *m, const float *x, const float *t, const float *w, int n, int iters, float lr, double *loss);
int learned_refresh(mgpu_ctx *ctx, const float *delta_x);
size_t learned_feature_blocks();
#define MGPU_HIP(ctx, call)                                                               \
  do {                                                                                    \
    hipError_t e_ = (call);                                                               \
    if (e_ != hipSuccess)                                                                 \
      return mgpu_fail((ctx), MGPU_E_HIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
  } while (0)

// phase timer: scoped accumulate into a field of ctx->t
struct PhaseTimer {
  mgpu_ctx *ctx;
  double *field;
  hipEvent_t e0 = nullptr, e1 = nullptr;
  PhaseTimer(mgpu_ctx *c, double *f);
  ~PhaseTimer();
};

// ---------------------------------------------------------------- FFT (fft.hip)
int fft_plan_init(mgpu_ctx *ctx, FFTPlan *plan, int n, int tcols = 16);
void fft_plan_free(FFTPlan *plan);
bool fft_tile16_fits(const mgpu_ctx *ctx, int n);
int fft_autotune(mgpu_ctx *ctx, FFTPlan *plan, float2 *buf, double ms[2]);
bool fft_size_supported(int n);
// in-place 3D transforms on a slot with the padded layout above
// in != nullptr: the real input is read from in (same padded layout) and buf only written
int fft_r2c_3d(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf, const float *in = nullptr);
// c2r; with out != nullptr the real result goes to a compact [N][N][N] array
int fft_c2r_3d(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf, float *out = nullptr);
// individual passes (used by fused paths and tests)
int fft_pass_z(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf, int forward, float *out = nullptr,
               const float *in = nullptr);
int fft_pass_col(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf, int axis, int forward);

// z c2r of the force's z component: X(k) -> i kz X(k) before the transform (in place)
int fft_pass_z_ikz(mgpu_ctx *ctx, const FFTPlan *plan, float2 *buf);

// ---------------------------------------------------------------- fused spectral passes (fused.hip)
enum { FUSED_GREEN_FORCE = 0, FUSED_GREEN_LAPLACE = 1, FUSED_GREEN_SMOOTH = 2 };
// inverse x pass of (a [+ b]) times the Green's function:
//   FORCE    o0 = X[i n_x G], o1 = X[G], G = S (-p0 / n^2)           (p0 = 1/(N^3 2pi/Box))
//   LAPLACE  o0 = X[S (-p0 / n^2)]                                   (p0 = DivideByLaplacian norm)
//   SMOOTH   o0 = X[S W(|n| p1) p0]                                  (p0 = 1/N^3, p1 = 2 pi R/Box)
int fused_green_x(mgpu_ctx *ctx, int mode, const float2 *a, const float2 *b, float2 *o0, float2 *o1, double p0,
                  double p1, int filter);
// inverse y pass: buf <- Y[buf], oy <- Y[i n_y buf]
int fused_dual_y(mgpu_ctx *ctx, float2 *buf, float2 *oy);
// forward x pass then times coupling n^2/(n^2 + massterm2)
int fused_epi_phi_x(mgpu_ctx *ctx, float2 *buf, double coupling, double massterm2);
// forward x pass of the density r2c, accumulating the binned P(k) power into pk[nbins]
int fused_pofk_x(mgpu_ctx *ctx, float2 *buf, double *pk);
// x pass pair (forward x transform, multiplier, inverse x transform(s)) in one kernel:
// fwd FUSED_FWD_DENSITY (in -> delta(k) to kout if not null, P(k) into pk if not null),
// FUSED_FWD_EPI (in = effective density; phi = coupling n^2/(n^2+m^2) X, S = dk + phi) or
// FUSED_FWD_PLAIN (S = dk + X); then mode FUSED_GREEN_* of S (p0, p1, filter as for
// fused_green_x) into o0 (and o1 for FUSED_GREEN_FORCE)
enum { FUSED_FWD_DENSITY = 0, FUSED_FWD_EPI = 1, FUSED_FWD_PLAIN = 2 };
int fused_x_pair(mgpu_ctx *ctx, int fwd, int mode, const float2 *in, float2 *kout, const float2 *dk, float2 *o0,
                 float2 *o1, double p0, double p1, int filter, double coupling, double massterm2, double *pk);
// z pass: c2r, screening against delta(x) (dx, real layout), r2c; in place on buf
int fused_z_screen(mgpu_ctx *ctx, int density_kind, float2 *buf, const float *dx, double p0, double p1);

// ---------------------------------------------------------------- tiled deposit (deposit.hip)
size_t deposit_scratch_bytes(int N, size_t np);
// particles -> delta(x) in real layout in d (sums in 64-bit fixed point in LDS)
// (with rsd != NULL: particles at the redshift-space records of k_rsd_coords instead)
int deposit_tiled(mgpu_ctx *ctx, float *d, const float4 *rsd = nullptr);
// Disp from the three force meshes by the last deposit's bins (sorted_valid must hold)
int gather_tiled(mgpu_ctx *ctx, const float *f0, const float *f1, const float *f2);

// ---------------------------------------------------------------- PM kernels (pm.hip)
int pm_deposit(mgpu_ctx *ctx);                         // particles -> delta(x) in slot0 (+ copy slot1)
int pm_pofk(mgpu_ctx *ctx, const float2 *dk, mgpu_pofk *out);
int pm_transform(mgpu_ctx *ctx, mgpu_pofk *out);  // r2c of delta(x) into slot 0 (+ P(k))
int pm_fifth_force(mgpu_ctx *ctx, const mgpu_mg *mg);
int pm_forces(mgpu_ctx *ctx);
// the whole of GetDisplacements with paired x passes (pm_step_fusable says when it applies)
int ic_nu_set(mgpu_ctx *ctx, double cdmfac, const mgpu_kslice *f);
int ic_nu_apply(mgpu_ctx *ctx);
int pm_rsd_pofk(mgpu_ctx *ctx, int axis, double vnorm, double s1, double s2, mgpu_rsd_pofk *out);
bool pm_step_fusable(const mgpu_ctx *ctx, const mgpu_mg *mg, const mgpu_pofk *out);
int pm_step_fused(mgpu_ctx *ctx, const mgpu_mg *mg, mgpu_pofk *out, double sumD[3]);
int pm_gather(mgpu_ctx *ctx, double sumD[3]);
int pm_kick(mgpu_ctx *ctx, double dda, double A, double s1, double s2, const double sumD[3], double sumV[3]);
int pm_drift(mgpu_ctx *ctx, double dyyy, double s1, double s2, const double sumV[3]);
int pm_kick_drift(mgpu_ctx *ctx, double dda, double A, double s1k, double s2k, const double sumD[3], double dyyy,
                  double s1d, double s2d, double sumV[3]);
int pm_reduce3(mgpu_ctx *ctx, int nblocks, double out[3], double *dst);

// ---------------------------------------------------------------- IC / LPT (ic.hip)
// seed_dev: nullptr allocates the seed table for the call; else the per-device scratch
int ic_displacement_fields(mgpu_ctx *ctx, unsigned long seed, int sphere_mode, int amp_fixed, int inverted,
                           const double *pk_n2, unsigned *seed_dev = nullptr);
// looks up this device's seed-table scratch (ctx->ic_scratch); called by mgpu_create
int ic_scratch_init(mgpu_ctx *ctx);
// the IC from delta(k) in cd on (2LPT source, lattice displacements), over ctx->bz realizations
int ic_lpt_from_delta(mgpu_ctx *ctx, float2 *cd);
// ic_displacement_fields for every member of a batch (m[0..count-1] in order), one launch per kernel
int ic_displacement_fields_batch(mgpu_ctx *const *m, int count, const unsigned long *seeds, int sphere_mode,
                                 int amp_fixed, const int *inverted, const double *pk_n2);
int ic_init_particles(mgpu_ctx *ctx, double Di, double Di2);
int ic_lpt_fields(mgpu_ctx *ctx, int slot, const double *g1_n2, const double *g2_n2);
int ic_lpt_fields_slices(mgpu_ctx *ctx, int slot, const mgpu_kslice *o1, const mgpu_kslice *o2);
// several slots at once (one fold reads the stored spectra once): q[2 s + o] = slice of (slots[s], order o+1)
int ic_lpt_fields_slices_n(mgpu_ctx *ctx, int ns, const int *slots, const mgpu_kslice *const *q);

static inline unsigned cdiv(size_t a, size_t b) { return (unsigned)((a + b - 1) / b); }
