// C ABI of libmgpu (include/mgpu.h): context management, particle I/O and the
// entry points that replace the reference's PM/COLA routines.
#include "mgpu_internal.h"

#include <cmath>
#include <cstdio>
#include <cstring>
#include <vector>

namespace {
thread_local std::string g_err;

int check_ctx(mgpu_ctx *ctx) { return ctx ? MGPU_OK : MGPU_E_INVALID; }

// one hold on a batch allocation dropped; the last frees it and the shared stream
void arena_release(mgpu_arena *a) {
  if (--a->refs > 0) return;
  if (a->base) (void)hipFree(a->base);
  if (a->stream) (void)hipStreamDestroy(a->stream);
  delete a;
}

int alloc(mgpu_ctx *ctx, void **p, size_t bytes, const char *what) {
  const size_t aligned = (bytes + 255) / 256 * 256;
  if (ctx->arena_cur) {
    // a batch member: carved from its share of the batch allocation
    if (aligned > (size_t)(ctx->arena_end - ctx->arena_cur)) {
      *p = nullptr;
      return mgpu_fail(ctx, MGPU_E_NOMEM, std::string("batch share exhausted at ") + what);
    }
    *p = ctx->arena_cur;
    ctx->arena_cur += aligned;
    return MGPU_OK;
  }
  if (ctx->arena_measure) *ctx->arena_measure += aligned;
  if (hipMalloc(p, bytes) != hipSuccess) {
    (void)hipGetLastError();
    *p = nullptr;
    return mgpu_fail(ctx, MGPU_E_NOMEM, std::string("device allocation failed: ") + what + " (" +
                                            std::to_string(bytes) + " bytes)");
  }
  return MGPU_OK;
}
}  // namespace

int mgpu_fail(mgpu_ctx *ctx, int code, const std::string &msg) {
  if (ctx) ctx->err = msg;
  g_err = msg;
  return code;
}

// Each scope records its own event pair so timers may nest (FFT time is also
// counted inside PtoMesh and Forces, as in the reference's timer.c).
PhaseTimer::PhaseTimer(mgpu_ctx *c, double *f) : ctx(c), field(f) {
  if (!ctx || !ctx->timers_on) return;
  if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
    (void)hipGetLastError();
    e0 = e1 = nullptr;
    return;
  }
  (void)hipEventRecord(e0, ctx->stream);
}
PhaseTimer::~PhaseTimer() {
  if (!e0) return;
  float ms = 0.f;
  if (hipEventRecord(e1, ctx->stream) == hipSuccess && hipEventSynchronize(e1) == hipSuccess &&
      hipEventElapsedTime(&ms, e0, e1) == hipSuccess)
    *field += ms;
  (void)hipGetLastError();
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
}

KScope::KScope(mgpu_ctx *c, int cat, double bytes) : ctx(c) {
  if (!ctx || !ctx->ktimers_on) return;
  r.cat = cat;
  r.bytes = bytes;
  if (hipEventCreate(&r.a) != hipSuccess || hipEventCreate(&r.b) != hipSuccess) {
    (void)hipGetLastError();
    return;
  }
  on = hipEventRecord(r.a, ctx->stream) == hipSuccess;
}
KScope::~KScope() {
  if (!on) return;
  if (hipEventRecord(r.b, ctx->stream) == hipSuccess) ctx->krec.push_back(r);
}

static int flush_ktimers(mgpu_ctx *ctx) {
  for (auto &r : ctx->krec) {
    float ms = 0.f;
    if (hipEventSynchronize(r.b) == hipSuccess && hipEventElapsedTime(&ms, r.a, r.b) == hipSuccess) {
      ctx->kcount[r.cat] += 1;
      ctx->kms[r.cat] += ms;
      ctx->kbytes[r.cat] += r.bytes;
    }
    (void)hipEventDestroy(r.a);
    (void)hipEventDestroy(r.b);
  }
  (void)hipGetLastError();
  ctx->krec.clear();
  return MGPU_OK;
}

extern "C" {

const char *mgpu_kernel_name(int k) {
  static const char *names[MGPU_K_COUNT] = {"deposit",    "mean",       "fft_r2c",  "fft_c2r", "fft_col_fwd",
                                            "fft_col_inv", "pofk",      "spectral", "real",    "gather",
                                            "kick",       "drift",      "lpt",      "readout", "fft_green",
                                            "fft_dual",   "fft_screen", "fft_epi",  "dep_sort", "fold",
                                            "fft_pot",    "fft_pk",     "fft_xdens", "fft_xforce"};
  return (k >= 0 && k < MGPU_K_COUNT) ? names[k] : "?";
}

int mgpu_enable_kernel_timing(mgpu_ctx *ctx, int on) {
  if (!ctx) return MGPU_E_INVALID;
  if (!on) flush_ktimers(ctx);
  ctx->ktimers_on = on != 0;
  return MGPU_OK;
}

int mgpu_kernel_timing(mgpu_ctx *ctx, int k, int *launches, double *ms, double *bytes, int reset) {
  if (!ctx || k < 0 || k >= MGPU_K_COUNT) return MGPU_E_INVALID;
  flush_ktimers(ctx);
  if (launches) *launches = ctx->kcount[k];
  if (ms) *ms = ctx->kms[k];
  if (bytes) *bytes = ctx->kbytes[k];
  if (reset) {
    ctx->kcount[k] = 0;
    ctx->kms[k] = 0;
    ctx->kbytes[k] = 0;
  }
  return MGPU_OK;
}

size_t mgpu_n2_table_len(int nmesh) { return (size_t)3 * (nmesh / 2) * (nmesh / 2) + 1; }

const char *mgpu_last_error(const void *ctx) {
  if (ctx) return ((const mgpu_ctx *)ctx)->err.c_str();
  return g_err.c_str();
}

int mgpu_device_info(char *name, size_t name_len, size_t *global_mem, int *cu_count, size_t *lds_per_block) {
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess || n == 0) {
    (void)hipGetLastError();
    g_err = "no HIP device visible";
    return MGPU_E_NODEVICE;
  }
  hipDeviceProp_t pr;
  if (hipGetDeviceProperties(&pr, 0) != hipSuccess) return MGPU_E_HIP;
  if (name && name_len) {
    std::string s = std::string(pr.name) + " (" + pr.gcnArchName + ")";
    std::strncpy(name, s.c_str(), name_len - 1);
    name[name_len - 1] = 0;
  }
  if (global_mem) *global_mem = pr.totalGlobalMem;
  if (cu_count) *cu_count = pr.multiProcessorCount;
  if (lds_per_block) *lds_per_block = pr.sharedMemPerBlock;
  return MGPU_OK;
}

// Device bytes of a context, by part (the plan mgpu_create allocates).
static void memory_plan(const mgpu_params *p, size_t part[MGPU_MEM_PARTS]) {
  const size_t N = p->nmesh, Np = p->nsample, np = Np * Np * Np;
  const size_t hp = ((N / 2 + 1) + 15) / 16 * 16, sb = N * N * hp * sizeof(float2);
  for (int i = 0; i < MGPU_MEM_PARTS; i++) part[i] = 0;
  part[MGPU_MEM_PARTICLES] = 5 * 3 * np * sizeof(float);
  part[MGPU_MEM_MESHES] = MGPU_NSLOTS * sb;
  part[MGPU_MEM_DEPOSIT] = deposit_scratch_bytes((int)N, np);
  if (p->scale_dependent) {
    part[MGPU_MEM_STORED] = 2 * sb;
    if (N % Np == 0 && fft_size_supported((int)Np)) {
      const size_t hps = ((Np / 2 + 1) + 15) / 16 * 16;
      part[MGPU_MEM_LATTICE] = 6 * Np * Np * hps * sizeof(float2);
    }
  }
  const size_t nb = (size_t)std::max(1, p->pofk_nbins);
  part[MGPU_MEM_TABLES] = (3 * std::max<size_t>((np + 255) / 256, N) + 3 * 1024) * sizeof(double) +
                          4 * nb * sizeof(double) + (N / 2 + 1) * 12 + 2 * N * sizeof(float2) +
                          mgpu_n2_table_len((int)N) * (2 * sizeof(double) + 2 * sizeof(float) + sizeof(int16_t));
  part[MGPU_MEM_RSD] = np * sizeof(float4);
}

int mgpu_memory_plan(const mgpu_params *p, size_t *parts, size_t *total) {
  if (!p || !fft_size_supported(p->nmesh) || p->nsample < 1 || p->nsample > 2048)
    return mgpu_fail(nullptr, MGPU_E_INVALID, "bad Nmesh or Nsample");
  size_t part[MGPU_MEM_PARTS];
  memory_plan(p, part);
  size_t t = 0;
  for (int i = 0; i < MGPU_MEM_PARTS; i++) {
    if (parts) parts[i] = part[i];
    if (i != MGPU_MEM_RSD) t += part[i];
  }
  if (total) *total = t;
  return MGPU_OK;
}

namespace {
// arena null: a context of its own (memory plan checked, own stream); else a batch
// member whose create-time buffers are carved from share (the arena's stream);
// measure not null: the aligned bytes of those buffers are added to *measure
int create_ctx(const mgpu_params *p, mgpu_ctx **out, mgpu_arena *arena, char *share, size_t *measure);
}  // namespace

int mgpu_create(const mgpu_params *p, mgpu_ctx **out) { return create_ctx(p, out, nullptr, nullptr, nullptr); }

namespace {
int create_ctx(const mgpu_params *p, mgpu_ctx **out, mgpu_arena *arena, char *share, size_t *measure) {
  if (!p || !out) return mgpu_fail(nullptr, MGPU_E_INVALID, "null argument");
  *out = nullptr;
  if (!fft_size_supported(p->nmesh))
    return mgpu_fail(nullptr, MGPU_E_INVALID, "Nmesh must be 2^a or 3*2^a in [8, 2048], got " + std::to_string(p->nmesh));
  if (p->nsample < 1 || p->nsample > 2048 || p->box <= 0)
    return mgpu_fail(nullptr, MGPU_E_INVALID, "bad Nsample or Box");
  if (p->model < MGPU_LCDM || p->model > MGPU_MBETA) return mgpu_fail(nullptr, MGPU_E_INVALID, "unknown model");
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return mgpu_fail(nullptr, MGPU_E_NODEVICE, "no HIP device visible");
  }
  if (!arena) {
    // the plan against the free device memory, before anything is allocated
    size_t part[MGPU_MEM_PARTS], need = 0, free_b = 0, total_b = 0;
    mgpu_memory_plan(p, part, &need);
    if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && need > free_b) {
      char msg[512];
      const double G = 1e9;
      std::snprintf(msg, sizeof msg,
                    "the run needs %.1f GB of device memory (particles %.1f, %d mesh slots %.1f, deposit %.1f, "
                    "stored spectra %.1f, lattice spectra %.1f, tables %.2f; RSD records %.1f more on the first "
                    "RSD call) but %.1f GB are free",
                    need / G, part[MGPU_MEM_PARTICLES] / G, MGPU_NSLOTS, part[MGPU_MEM_MESHES] / G,
                    part[MGPU_MEM_DEPOSIT] / G, part[MGPU_MEM_STORED] / G, part[MGPU_MEM_LATTICE] / G,
                    part[MGPU_MEM_TABLES] / G, part[MGPU_MEM_RSD] / G, free_b / G);
      return mgpu_fail(nullptr, MGPU_E_NOMEM, msg);
    }
    (void)hipGetLastError();
  }
  mgpu_ctx *ctx = new mgpu_ctx();
  ctx->p = *p;
  ctx->N = p->nmesh;
  ctx->Np = p->nsample;
  ctx->np = (size_t)p->nsample * p->nsample * p->nsample;
  int rc;
  auto fail = [&](int code) {
    g_err = ctx->err;
    mgpu_destroy(ctx);
    return code;
  };
  if (arena) {
    ctx->arena = arena;
    ctx->arena_cur = share;
    ctx->arena_end = share + arena->stride;
    ctx->zs = arena->stride;
    ctx->stream = arena->stream;
    arena->refs++;
  }
  ctx->arena_measure = measure;
  if (hipGetDevice(&ctx->device) != hipSuccess) return fail(mgpu_fail(ctx, MGPU_E_HIP, "hipGetDevice failed"));
  if ((rc = ic_scratch_init(ctx))) return fail(rc);
  hipDeviceProp_t pr;
  if (hipGetDeviceProperties(&pr, ctx->device) == hipSuccess) ctx->lds_max = pr.sharedMemPerBlock;
  if (!ctx->stream && hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking) != hipSuccess)
    return fail(mgpu_fail(ctx, MGPU_E_HIP, "stream creation failed"));
  if ((rc = fft_plan_init(ctx, &ctx->fft, ctx->N, MGPU_FFT_TCOLS))) return fail(rc);
  ctx->HP = ctx->fft.hp;
  ctx->slot_elems = (size_t)ctx->N * ctx->N * ctx->HP;
  const size_t pb = 3 * ctx->np * sizeof(float);
  if ((rc = alloc(ctx, (void **)&ctx->pos, pb, "positions"))) return fail(rc);
  if ((rc = alloc(ctx, (void **)&ctx->vel, pb, "velocities"))) return fail(rc);
  if ((rc = alloc(ctx, (void **)&ctx->disp, pb, "Disp"))) return fail(rc);
  for (int s = 0; s < 2; s++)
    if ((rc = alloc(ctx, (void **)&ctx->lpt[s], pb, "LPT slot"))) return fail(rc);
  const size_t sb = ctx->slot_elems * sizeof(float2);
  for (int s = 0; s < MGPU_NSLOTS; s++)
    if ((rc = alloc(ctx, (void **)&ctx->slot[s], sb, "mesh slot"))) return fail(rc);
  ctx->dep_scratch_bytes = deposit_scratch_bytes(ctx->N, ctx->np);
  if ((rc = alloc(ctx, &ctx->dep_scratch, ctx->dep_scratch_bytes, "deposit scratch"))) return fail(rc);
  if (p->scale_dependent) {
    for (int s = 0; s < 2; s++)
      if ((rc = alloc(ctx, (void **)&ctx->cdelta[s], sb, "stored delta(k)"))) return fail(rc);
    // LPT fields by the exact lattice fold when the mesh is a multiple of the lattice
    if (ctx->N % ctx->Np == 0 && fft_size_supported(ctx->Np)) {
      if ((rc = fft_plan_init(ctx, &ctx->fft_lattice, ctx->Np, MGPU_FFT_TCOLS))) return fail(rc);
      ctx->fft_lattice.kcat = MGPU_K_LPT;
      const size_t nl = (size_t)ctx->Np * ctx->Np * ctx->fft_lattice.hp;
      if ((rc = alloc(ctx, (void **)&ctx->lattice_slot, 6 * nl * sizeof(float2), "lattice spectra"))) return fail(rc);
    }
  }
  ctx->partial_len = 3 * (size_t)std::max<size_t>(cdiv(ctx->np, 256), (size_t)ctx->N) + 3 * 1024;
  if ((rc = alloc(ctx, (void **)&ctx->partial, ctx->partial_len * sizeof(double), "reduction scratch"))) return fail(rc);
  if ((rc = alloc(ctx, (void **)&ctx->dsum, 8 * sizeof(double), "sums"))) return fail(rc);
  const int nb = std::max(1, p->pofk_nbins);
  if ((rc = alloc(ctx, (void **)&ctx->pk, 4 * (size_t)nb * sizeof(double), "P(k) bins"))) return fail(rc);
  {
    // compute_power_spectrum's CIC deconvolution 1/(W_x W_y W_z)^4 with the 1/N^6 of the
    // transform split over the three axes (src/compute_pofk.c:128-134)
    std::vector<double> winv(ctx->N / 2 + 1);
    const double n6 = 1.0 / std::pow((double)ctx->N, 6);
    for (int d = 0; d <= ctx->N / 2; d++) {
      const double g = d == 0 ? 1.0 : std::sin((MGPU_PI * d) / (double)ctx->N) / ((MGPU_PI * d) / (double)ctx->N);
      winv[d] = 1.0 / std::pow(g, 4.0) * std::cbrt(n6);
    }
    if ((rc = alloc(ctx, (void **)&ctx->winv, winv.size() * sizeof(double), "P(k) window"))) return fail(rc);
    if (hipMemcpy(ctx->winv, winv.data(), winv.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess)
      return fail(mgpu_fail(ctx, MGPU_E_HIP, "window upload failed"));
    std::vector<float> winvf(winv.begin(), winv.end());
    if ((rc = alloc(ctx, (void **)&ctx->winvf, winvf.size() * sizeof(float), "P(k) window"))) return fail(rc);
    if (hipMemcpy(ctx->winvf, winvf.data(), winvf.size() * sizeof(float), hipMemcpyHostToDevice) != hipSuccess)
      return fail(mgpu_fail(ctx, MGPU_E_HIP, "window upload failed"));
  }
  ctx->n2len = mgpu_n2_table_len(ctx->N);
  if (p->pofk_nbins > 0 && p->pofk_nbins <= 32767) {
    // P(k) bin of every |n|^2, by compute_power_spectrum's formula in FP64
    // (src/compute_pofk.c:112-126); -1 outside the bins
    std::vector<int16_t> tab(ctx->n2len);
    const double kmin = p->pofk_kmin, kmax = p->pofk_kmax, lnrange = p->pofk_bintype == 1 ? std::log(kmax / kmin) : 1.0;
    for (size_t n2 = 0; n2 < ctx->n2len; n2++) {
      const double kmag = std::sqrt((double)n2);
      int nk;
      if (p->pofk_bintype == 0) nk = (int)((kmag - kmin) / (kmax - kmin) * p->pofk_nbins + 0.5);
      else nk = kmag <= 0.0 ? -1 : (int)(std::log(kmag / kmin) / lnrange * p->pofk_nbins + 0.5);
      tab[n2] = (int16_t)(nk < 0 || nk >= p->pofk_nbins ? -1 : nk);
    }
    if ((rc = alloc(ctx, (void **)&ctx->pk_bin, tab.size() * sizeof(int16_t), "P(k) bin table"))) return fail(rc);
    if (hipMemcpy(ctx->pk_bin, tab.data(), tab.size() * sizeof(int16_t), hipMemcpyHostToDevice) != hipSuccess)
      return fail(mgpu_fail(ctx, MGPU_E_HIP, "bin table upload failed"));
  }
  for (int s = 0; s < 2; s++) {
    if ((rc = alloc(ctx, (void **)&ctx->n2tab[s], ctx->n2len * sizeof(double), "n2 table"))) return fail(rc);
    if ((rc = alloc(ctx, (void **)&ctx->n2tabf[s], ctx->n2len * sizeof(float), "n2 table"))) return fail(rc);
  }
  if (hipEventCreate(&ctx->ev0) != hipSuccess || hipEventCreate(&ctx->ev1) != hipSuccess)
    return fail(mgpu_fail(ctx, MGPU_E_HIP, "event creation failed"));
  // defined contents everywhere
  if (hipMemsetAsync(ctx->vel, 0, pb, ctx->stream) != hipSuccess ||
      hipMemsetAsync(ctx->dsum, 0, 8 * sizeof(double), ctx->stream) != hipSuccess ||
      hipMemsetAsync(ctx->lpt[0], 0, pb, ctx->stream) != hipSuccess ||
      hipMemsetAsync(ctx->lpt[1], 0, pb, ctx->stream) != hipSuccess ||
      hipMemsetAsync(ctx->disp, 0, pb, ctx->stream) != hipSuccess)
    return fail(mgpu_fail(ctx, MGPU_E_HIP, "memset failed"));
  // meshes too: the padding entries of each row are never written by the passes
  for (int s = 0; s < MGPU_NSLOTS; s++)
    if (hipMemsetAsync(ctx->slot[s], 0, sb, ctx->stream) != hipSuccess)
      return fail(mgpu_fail(ctx, MGPU_E_HIP, "memset failed"));
  if ((ctx->lattice_slot && hipMemsetAsync(ctx->lattice_slot, 0,
                                           6 * (size_t)ctx->Np * ctx->Np * ctx->fft_lattice.hp * sizeof(float2),
                                           ctx->stream) != hipSuccess))
    return fail(mgpu_fail(ctx, MGPU_E_HIP, "memset failed"));
  if (hipStreamSynchronize(ctx->stream) != hipSuccess) return fail(mgpu_fail(ctx, MGPU_E_HIP, "sync failed"));
  // later allocations (RSD records, learned tables, slices) are the member's own
  ctx->arena_cur = ctx->arena_end = nullptr;
  ctx->arena_measure = nullptr;
  *out = ctx;
  return MGPU_OK;
}
}  // namespace

void mgpu_destroy(mgpu_ctx *ctx) {
  if (!ctx) return;
  if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
  mgpu_arena *arena = ctx->arena;
  const char *a0 = arena ? arena->base : nullptr, *a1 = arena ? arena->base + arena->stride * arena->count : nullptr;
  flush_ktimers(ctx);
  if (ctx->step_exec) (void)hipGraphExecDestroy(ctx->step_exec);
  void *ptrs[] = {ctx->pos,      ctx->vel,       ctx->disp,      ctx->lpt[0],    ctx->lpt[1],
                  ctx->slot[0],  ctx->slot[1],   ctx->slot[2],   ctx->slot[3],   ctx->cdelta[0],
                  ctx->cdelta[1], ctx->partial,  ctx->dsum,      ctx->pk,        ctx->n2tab[0],
                  ctx->n2tab[1], ctx->n2tabf[0], ctx->n2tabf[1], ctx->lattice_slot, ctx->slice_buf,
                  ctx->dep_scratch, ctx->winv,    ctx->winvf,     ctx->pk_bin,    ctx->rsd_crd,
                  ctx->nu_slice, ctx->nu_fac,  ctx->feat_part, ctx->feat,     ctx->green_mult};
  for (void *q : ptrs)
    if (q && !((const char *)q >= a0 && (const char *)q < a1)) (void)hipFree(q);
  if (ctx->slice_host) (void)hipHostFree(ctx->slice_host);
  if (ctx->slice_ev) (void)hipEventDestroy(ctx->slice_ev);
  fft_plan_free(&ctx->fft);
  fft_plan_free(&ctx->fft_lattice);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  if (arena) {
    // the last member frees the batch allocation and the shared stream
    arena_release(arena);
  } else if (ctx->stream) {
    (void)hipStreamDestroy(ctx->stream);
  }
  delete ctx;
}

int mgpu_sync(mgpu_ctx *ctx) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

int mgpu_set_particles(mgpu_ctx *ctx, const float *pos, const float *vel) {
  if (ctx) ctx->sorted_valid = false;
  if (ctx && vel) ctx->vsum_valid = false;
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  const size_t b = 3 * ctx->np * sizeof(float);
  if (pos) MGPU_HIP(ctx, hipMemcpyAsync(ctx->pos, pos, b, hipMemcpyHostToDevice, ctx->stream));
  if (vel) MGPU_HIP(ctx, hipMemcpyAsync(ctx->vel, vel, b, hipMemcpyHostToDevice, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

int mgpu_get_particles(mgpu_ctx *ctx, float *pos, float *vel) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  const size_t b = 3 * ctx->np * sizeof(float);
  if (pos) MGPU_HIP(ctx, hipMemcpyAsync(pos, ctx->pos, b, hipMemcpyDeviceToHost, ctx->stream));
  if (vel) MGPU_HIP(ctx, hipMemcpyAsync(vel, ctx->vel, b, hipMemcpyDeviceToHost, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

int mgpu_set_lpt(mgpu_ctx *ctx, int slot, const float *d) {
  if (check_ctx(ctx) || slot < 0 || slot > 1 || !d) return mgpu_fail(ctx, MGPU_E_INVALID, "bad LPT slot");
  MGPU_HIP(ctx, hipMemcpyAsync(ctx->lpt[slot], d, 3 * ctx->np * sizeof(float), hipMemcpyHostToDevice, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

int mgpu_get_lpt(mgpu_ctx *ctx, int slot, float *d) {
  if (check_ctx(ctx) || slot < 0 || slot > 1 || !d) return mgpu_fail(ctx, MGPU_E_INVALID, "bad LPT slot");
  MGPU_HIP(ctx, hipMemcpyAsync(d, ctx->lpt[slot], 3 * ctx->np * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

int mgpu_get_disp(mgpu_ctx *ctx, float *d) {
  if (check_ctx(ctx) || !d) return MGPU_E_INVALID;
  MGPU_HIP(ctx, hipMemcpyAsync(d, ctx->disp, 3 * ctx->np * sizeof(float), hipMemcpyDeviceToHost, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

int mgpu_pm_deposit(mgpu_ctx *ctx) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  int rc = pm_deposit(ctx);
  if (rc) return rc;
  ctx->have_deposit = true;
  ctx->have_density = false;
  ctx->have_forces = false;
  ctx->have_phik = false;
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

int mgpu_pm_transform(mgpu_ctx *ctx, mgpu_pofk *pk) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  if (!ctx->have_deposit) return mgpu_fail(ctx, MGPU_E_STATE, "no deposited density: call mgpu_pm_deposit first");
  // delta(x) stays in slot dx_slot when it is needed later (MG screening)
  int rc = pm_transform(ctx, pk);
  if (rc) return rc;
  ctx->have_deposit = false;
  ctx->have_density = true;
  return MGPU_OK;
}

int mgpu_pm_density(mgpu_ctx *ctx, mgpu_pofk *pk) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  int rc;
  if ((rc = pm_deposit(ctx))) return rc;
  ctx->have_deposit = true;
  ctx->have_forces = false;
  ctx->have_phik = false;
  if ((rc = mgpu_pm_transform(ctx, pk))) return rc;
  // the neutrino term follows the CDM P(k) (src/auxPM.c:367-420)
  return ctx->nu_on ? ic_nu_apply(ctx) : MGPU_OK;
}

int mgpu_set_neutrinos(mgpu_ctx *ctx, double cdmfac, const mgpu_kslice *fac) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  return ic_nu_set(ctx, cdmfac, fac);
}

int mgpu_neutrino_response(mgpu_ctx *ctx) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  if (!ctx->have_density) return mgpu_fail(ctx, MGPU_E_STATE, "no density(k): call mgpu_pm_density first");
  return ic_nu_apply(ctx);
}

int mgpu_growth_tables(const mgpu_growth_model *m, double x0, double dx, int na, const double *lnk, int nk,
                       double *out) {
  if (!m || !lnk || !out || na < 1 || nk < 1 || !(dx > 0) || !(m->omega > 0))
    return mgpu_fail(nullptr, MGPU_E_INVALID, "growth tables: bad argument");
  if (m->model != MGPU_LCDM && m->model != MGPU_FOFR && m->model != MGPU_MBETA)
    return mgpu_fail(nullptr, MGPU_E_INVALID, "growth tables: scale-dependent models are LCDM, f(R) and m/beta");
  if (m->mg_active && m->model == MGPU_FOFR && !(m->fofr0 > 0))
    return mgpu_fail(nullptr, MGPU_E_INVALID, "growth tables: fofr0 must be positive");
  if (m->mg_active && m->model == MGPU_MBETA && !(m->range_symm > 0))
    return mgpu_fail(nullptr, MGPU_E_INVALID, "growth tables: range_symm must be positive");
  if (m->omega_nu > 1e-6) {
    const mgpu_spline2d *t[3] = {&m->t_nu, &m->t_cdm, &m->t_b};
    for (const mgpu_spline2d *s : t)
      if (s->nx < 2 || s->ny < 2 || !s->x || !s->y || !s->z || !s->zx || !s->zy || !s->zxy)
        return mgpu_fail(nullptr, MGPU_E_INVALID, "growth tables: neutrino factors need the three transfer splines");
  }
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0) {
    (void)hipGetLastError();
    return mgpu_fail(nullptr, MGPU_E_NODEVICE, "no HIP device visible");
  }
  const int rc = growth_tables(m, x0, dx, na, lnk, nk, out);
  if (rc == MGPU_E_INVALID) return mgpu_fail(nullptr, rc, "growth tables: an ODE failed (NaN or 10^6 steps)");
  if (rc) return mgpu_fail(nullptr, rc, "growth tables: HIP allocation or copy failed");
  return MGPU_OK;
}

int mgpu_mlp_create(unsigned long long seed, mgpu_mlp **out) {
  if (!out) return mgpu_fail(nullptr, MGPU_E_INVALID, "null argument");
  *out = nullptr;
  return mlp_create(seed, out);
}

void mgpu_mlp_destroy(mgpu_mlp *m) { mlp_destroy(m); }

int mgpu_mlp_get_params(mgpu_mlp *m, float *params) {
  if (!m || !params) return mgpu_fail(nullptr, MGPU_E_INVALID, "null argument");
  return mlp_params(m, params, nullptr);
}

int mgpu_mlp_set_params(mgpu_mlp *m, const float *params) {
  if (!m || !params) return mgpu_fail(nullptr, MGPU_E_INVALID, "null argument");
  return mlp_params(m, nullptr, params);
}

int mgpu_mlp_forward(mgpu_mlp *m, const float *x, int n, float *y) {
  if (!m || !x || !y || n < 1) return mgpu_fail(nullptr, MGPU_E_INVALID, "MLP forward: bad argument");
  return mlp_forward(m, x, n, y);
}

int mgpu_mlp_gradient(mgpu_mlp *m, const float *x, const float *t, const float *w, int n, float *grad,
                      double *loss) {
  if (!m || !x || !t || !w || n < 1) return mgpu_fail(nullptr, MGPU_E_INVALID, "MLP gradient: bad argument");
  return mlp_gradient(m, x, t, w, n, grad, loss);
}

int mgpu_mlp_train(mgpu_mlp *m, const float *x, const float *t, const float *w, int n, int iters, float lr,
                   double *loss) {
  if (!m || !x || !t || !w || n < 1 || iters < 0 || !(lr > 0))
    return mgpu_fail(nullptr, MGPU_E_INVALID, "MLP training: bad argument");
  return mlp_train(m, x, t, w, n, iters, lr, loss);
}

int mgpu_set_learned_multiplier(mgpu_ctx *ctx, mgpu_mlp *m, double a, double theta) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  if (a < 0) {
    ctx->learn_on = false;
    ctx->mlp = nullptr;
    return MGPU_OK;
  }
  int rc;
  if (!ctx->feat && ((rc = alloc(ctx, (void **)&ctx->feat_part, learned_feature_blocks() * sizeof(double), "feature")) ||
                     (rc = alloc(ctx, (void **)&ctx->feat, sizeof(double), "feature")) ||
                     (rc = alloc(ctx, (void **)&ctx->green_mult, ctx->n2len * sizeof(float), "multiplier table"))))
    return rc;
  ctx->learn_on = true;
  ctx->mlp = m;
  ctx->learn_a = a;
  ctx->learn_theta = theta;
  return MGPU_OK;
}

int mgpu_density_feature(mgpu_ctx *ctx, double *s) {
  if (check_ctx(ctx) || !s) return MGPU_E_INVALID;
  if (!ctx->feat) return mgpu_fail(ctx, MGPU_E_STATE, "no density feature: call mgpu_set_learned_multiplier first");
  MGPU_HIP(ctx, hipMemcpyAsync(s, ctx->feat, sizeof(double), hipMemcpyDeviceToHost, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

static FFTPlan *plan_of(mgpu_ctx *ctx, int which) {
  if (which == 0) return &ctx->fft;
  if (which == 1 && ctx->fft_lattice.n > 0) return &ctx->fft_lattice;
  return nullptr;
}

int mgpu_set_fft_tile(mgpu_ctx *ctx, int which, int tcols) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  FFTPlan *pl = plan_of(ctx, which);
  if (!pl) return mgpu_fail(ctx, MGPU_E_INVALID, "no such FFT plan");
  if (tcols != 8 && tcols != 16) return mgpu_fail(ctx, MGPU_E_INVALID, "tile width must be 8 or 16");
  if (tcols == 16 && !fft_tile16_fits(ctx, pl->n))
    return mgpu_fail(ctx, MGPU_E_INVALID, "16-column tiles do not fit a workgroup at this size");
  pl->tcols = tcols;
  return MGPU_OK;
}

int mgpu_get_fft_tile(mgpu_ctx *ctx, int which, int *tcols) {
  if (check_ctx(ctx) || !tcols) return MGPU_E_INVALID;
  FFTPlan *pl = plan_of(ctx, which);
  if (!pl) return mgpu_fail(ctx, MGPU_E_INVALID, "no such FFT plan");
  *tcols = pl->tcols;
  return MGPU_OK;
}

int mgpu_autotune_fft(mgpu_ctx *ctx, double ms[4]) {
  if (check_ctx(ctx) || !ms) return MGPU_E_INVALID;
  ms[0] = ms[1] = ms[2] = ms[3] = -1.0;
  int rc = fft_autotune(ctx, &ctx->fft, ctx->slot[3], ms);
  if (!rc && ctx->fft_lattice.n > 0 && ctx->lattice_slot) rc = fft_autotune(ctx, &ctx->fft_lattice, ctx->lattice_slot, ms + 2);
  return rc;
}

int mgpu_power_spectrum(mgpu_ctx *ctx, mgpu_pofk *pk) {
  if (check_ctx(ctx) || !pk) return MGPU_E_INVALID;
  if (!ctx->have_density) return mgpu_fail(ctx, MGPU_E_STATE, "no density(k): call mgpu_pm_density first");
  return pm_pofk(ctx, ctx->slot[0], pk);
}

int mgpu_fifth_force(mgpu_ctx *ctx, const mgpu_mg *mg) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  if (!ctx->have_density) return mgpu_fail(ctx, MGPU_E_STATE, "no density(k): call mgpu_pm_density first");
  return pm_fifth_force(ctx, mg);
}

int mgpu_forces(mgpu_ctx *ctx) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  if (!ctx->have_density) return mgpu_fail(ctx, MGPU_E_STATE, "no density(k): call mgpu_pm_density first");
  int rc = pm_forces(ctx);
  if (rc) return rc;
  ctx->have_forces = true;
  return MGPU_OK;
}

int mgpu_mesh_to_particles(mgpu_ctx *ctx, double sumD[3]) {
  if (check_ctx(ctx) || !sumD) return MGPU_E_INVALID;
  if (!ctx->have_forces) return mgpu_fail(ctx, MGPU_E_STATE, "no force meshes: call mgpu_forces first");
  int rc = pm_gather(ctx, sumD);
  if (rc) return rc;
  ctx->have_disp = true;
  return MGPU_OK;
}

int mgpu_get_displacements(mgpu_ctx *ctx, const mgpu_mg *mg, mgpu_pofk *pk, double sumD[3]) {
  int rc;
  if (check_ctx(ctx) || !sumD) return MGPU_E_INVALID;
  if (ctx->p.mg_active && ctx->p.model != MGPU_LCDM && !mg)
    return mgpu_fail(ctx, MGPU_E_INVALID, "modified gravity active but no mgpu_mg coefficients given");
  if (pm_step_fusable(ctx, mg, pk)) {
    ctx->have_deposit = ctx->have_density = ctx->have_forces = false;
    ctx->have_phik = false;
    if ((rc = pm_step_fused(ctx, mg, pk, sumD))) return rc;
    ctx->have_density = ctx->have_forces = ctx->have_disp = true;
    return MGPU_OK;
  }
  if ((rc = mgpu_pm_density(ctx, pk))) return rc;
  if (ctx->p.mg_active && (rc = mgpu_fifth_force(ctx, mg))) return rc;
  if ((rc = mgpu_forces(ctx))) return rc;
  return mgpu_mesh_to_particles(ctx, sumD);
}

int mgpu_rsd_power_spectrum(mgpu_ctx *ctx, int axis, double vnorm, double s1, double s2, mgpu_rsd_pofk *out) {
  if (check_ctx(ctx) || !out) return MGPU_E_INVALID;
  return pm_rsd_pofk(ctx, axis, vnorm, s1, s2, out);
}

int mgpu_kick(mgpu_ctx *ctx, double dda, double A, double s1, double s2, const double sumD[3], double sumV[3]) {
  if (check_ctx(ctx) || !sumD || !sumV) return MGPU_E_INVALID;
  if (!ctx->have_disp) return mgpu_fail(ctx, MGPU_E_STATE, "no Disp: call mgpu_mesh_to_particles first");
  return pm_kick(ctx, dda, A, s1, s2, sumD, sumV);
}

int mgpu_drift(mgpu_ctx *ctx, double dyyy, double s1, double s2, const double sumV[3]) {
  if (check_ctx(ctx) || !sumV) return MGPU_E_INVALID;
  return pm_drift(ctx, dyyy, s1, s2, sumV);
}

int mgpu_kick_drift(mgpu_ctx *ctx, double dda, double A, double s1k, double s2k, const double sumD[3], double dyyy,
                    double s1d, double s2d, double sumV[3]) {
  if (check_ctx(ctx) || !sumD || !sumV) return MGPU_E_INVALID;
  if (!ctx->have_disp) return mgpu_fail(ctx, MGPU_E_STATE, "no Disp: call mgpu_mesh_to_particles first");
  const int rc = pm_kick_drift(ctx, dda, A, s1k, s2k, sumD, dyyy, s1d, s2d, sumV);
  ctx->have_disp = false;  // Disp was not reduced by sumD
  return rc;
}

// The step body of mgpu_step: the fused GetDisplacements (its Disp sums left in
// dsum[0..2]) and Kick + Drift reading the means from dsum; no host wait.
static int device_step_body(mgpu_ctx *ctx, const mgpu_mg *mg, double dda, double A, double s1k, double s2k,
                            double dyyy, double s1d, double s2d) {
  ctx->have_deposit = ctx->have_density = ctx->have_forces = false;
  ctx->have_phik = false;
  int rc = pm_step_fused(ctx, mg, nullptr, nullptr);
  if (rc) return rc;
  return pm_kick_drift(ctx, dda, A, s1k, s2k, nullptr, dyyy, s1d, s2d, nullptr);
}

namespace {
// what mgpu_step needs of a context (MGPU_OK or the refusal)
int step_check(mgpu_ctx *ctx, const mgpu_mg *mg) {
  if (ctx->p.mg_active && ctx->p.model != MGPU_LCDM && !mg)
    return mgpu_fail(ctx, MGPU_E_INVALID, "modified gravity active but no mgpu_mg coefficients given");
  if (ctx->p.scale_dependent)
    return mgpu_fail(ctx, MGPU_E_STATE,
                     "scale-dependent runs refill the LPT fields between the force and the kick: use the separate calls");
  if (!pm_step_fusable(ctx, mg, nullptr) || ctx->learn_on || ctx->nu_on)
    return mgpu_fail(ctx, MGPU_E_STATE, "this model's step is not one fused sequence: use the separate calls");
  if (ctx->timers_on || ctx->ktimers_on)
    return mgpu_fail(ctx, MGPU_E_STATE, "timers wait for the device inside a step: disable them for mgpu_step");
  if (!ctx->vsum_valid)
    return mgpu_fail(ctx, MGPU_E_STATE, "no velocity sums on the device: the first step takes mgpu_kick or "
                                        "mgpu_kick_drift");
  return MGPU_OK;
}

// the step body on ctx's stream, eagerly or (graph, after one eager step) captured
// into ctx's executable graph, updated in place, and launched once
int run_step(mgpu_ctx *ctx, const mgpu_mg *mg, double dda, double A, double s1k, double s2k, double dyyy,
             double s1d, double s2d, int graph) {
  int rc;
  if (graph && ctx->step_warm) {
    hipGraph_t g = nullptr;
    MGPU_HIP(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeRelaxed));
    rc = device_step_body(ctx, mg, dda, A, s1k, s2k, dyyy, s1d, s2d);
    const hipError_t e = hipStreamEndCapture(ctx->stream, &g);
    if (rc || e != hipSuccess || !g) {
      if (g) (void)hipGraphDestroy(g);
      (void)hipGetLastError();
      return rc ? rc : mgpu_fail(ctx, MGPU_E_HIP, std::string("step capture failed: ") + hipGetErrorString(e));
    }
    if (ctx->step_exec) {
      // same kernels and shapes every step: only the scalar arguments change
      hipGraphNode_t en = nullptr;
      hipGraphExecUpdateResult ur = hipGraphExecUpdateSuccess;
      if (hipGraphExecUpdate(ctx->step_exec, g, &en, &ur) != hipSuccess || ur != hipGraphExecUpdateSuccess) {
        (void)hipGetLastError();
        (void)hipGraphExecDestroy(ctx->step_exec);
        ctx->step_exec = nullptr;
      }
    }
    if (!ctx->step_exec && hipGraphInstantiate(&ctx->step_exec, g, nullptr, nullptr, 0) != hipSuccess) {
      ctx->step_exec = nullptr;
      (void)hipGraphDestroy(g);
      return mgpu_fail(ctx, MGPU_E_HIP, "step graph instantiation failed");
    }
    (void)hipGraphDestroy(g);
    MGPU_HIP(ctx, hipGraphLaunch(ctx->step_exec, ctx->stream));
  } else {
    if ((rc = device_step_body(ctx, mg, dda, A, s1k, s2k, dyyy, s1d, s2d))) return rc;
    ctx->step_warm = true;
  }
  ctx->have_density = ctx->have_forces = true;
  ctx->have_disp = false;  // Disp was not reduced by the mean
  return MGPU_OK;
}
}  // namespace

int mgpu_step(mgpu_ctx *ctx, const mgpu_mg *mg, double dda, double A, double s1k, double s2k, double dyyy,
              double s1d, double s2d, int graph) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  int rc = step_check(ctx, mg);
  if (rc) return rc;
  return run_step(ctx, mg, dda, A, s1k, s2k, dyyy, s1d, s2d, graph);
}

int mgpu_batch_create(const mgpu_params *p, int count, mgpu_ctx **members) {
  if (!p || !members || count < 1) return mgpu_fail(nullptr, MGPU_E_INVALID, "bad batch arguments");
  for (int r = 0; r < count; r++) members[r] = nullptr;
  // one context made and destroyed to measure a member's share
  size_t share = 0;
  {
    mgpu_ctx *probe = nullptr;
    int rc = create_ctx(p, &probe, nullptr, nullptr, &share);
    if (rc) return rc;
    mgpu_destroy(probe);
  }
  share = (share + 255) / 256 * 256;
  if (share > SIZE_MAX / (size_t)count) return mgpu_fail(nullptr, MGPU_E_NOMEM, "batch too large");
  const size_t total = share * (size_t)count;
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) == hipSuccess && total > free_b)
    return mgpu_fail(nullptr, MGPU_E_NOMEM, "the batch needs " + std::to_string(total) + " bytes of device memory but " +
                                                std::to_string(free_b) + " are free");
  (void)hipGetLastError();
  mgpu_arena *arena = new mgpu_arena();
  arena->stride = share;
  arena->count = count;
  arena->refs = 1;  // held by this call until the members exist
  if (hipMalloc((void **)&arena->base, total) != hipSuccess) {
    (void)hipGetLastError();
    arena->base = nullptr;
    arena_release(arena);
    return mgpu_fail(nullptr, MGPU_E_NOMEM, "batch allocation failed (" + std::to_string(total) + " bytes)");
  }
  if (hipStreamCreateWithFlags(&arena->stream, hipStreamNonBlocking) != hipSuccess) {
    (void)hipGetLastError();
    arena->stream = nullptr;
    arena_release(arena);
    return mgpu_fail(nullptr, MGPU_E_HIP, "stream creation failed");
  }
  int rc = MGPU_OK;
  for (int r = 0; r < count && !rc; r++) rc = create_ctx(p, &members[r], arena, arena->base + share * r, nullptr);
  if (!rc) {
    // every member's buffers at its share's offset: realization r = realization 0 + r * share
    mgpu_ctx *m0 = members[0];
    for (int r = 1; r < count && !rc; r++) {
      const mgpu_ctx *m = members[r];
      const void *a[] = {m->pos, m->vel, m->disp, m->lpt[0], m->lpt[1], m->slot[0], m->slot[1], m->slot[2], m->slot[3],
                         m->partial, m->dsum, m->dep_scratch};
      const void *b[] = {m0->pos, m0->vel, m0->disp, m0->lpt[0], m0->lpt[1], m0->slot[0], m0->slot[1], m0->slot[2],
                         m0->slot[3], m0->partial, m0->dsum, m0->dep_scratch};
      for (size_t q = 0; q < sizeof a / sizeof a[0]; q++)
        if ((const char *)a[q] != (const char *)b[q] + share * r)
          rc = mgpu_fail(nullptr, MGPU_E_HIP, "batch members' buffers are not at a fixed stride");
      if (m->fft.tcols != m0->fft.tcols) rc = mgpu_fail(nullptr, MGPU_E_HIP, "batch members' FFT tiles differ");
    }
  }
  if (rc) {
    const std::string e = g_err;
    for (int r = 0; r < count; r++) {
      mgpu_destroy(members[r]);
      members[r] = nullptr;
    }
    arena_release(arena);
    g_err = e;
    return rc;
  }
  arena_release(arena);
  return MGPU_OK;
}

namespace {
// all the contexts of one mgpu_batch_create, in order
int batch_members(mgpu_ctx *const *members, int count) {
  if (!members || count < 1 || !members[0]) return mgpu_fail(nullptr, MGPU_E_INVALID, "bad batch arguments");
  mgpu_ctx *m0 = members[0];
  mgpu_arena *arena = m0->arena;
  for (int r = 0; arena && arena->count == count && r < count; r++) {
    const mgpu_ctx *m = members[r];
    if (!m || m->arena != arena || m->pos != (float *)((char *)m0->pos + arena->stride * r)) arena = nullptr;
  }
  if (!arena || arena->count != count)
    return mgpu_fail(m0, MGPU_E_INVALID, "the members must be all the contexts of one mgpu_batch_create, in order");
  return MGPU_OK;
}
}  // namespace

int mgpu_batch_ic_displacement_fields(mgpu_ctx *const *members, int count, const unsigned long *seeds,
                                      int sphere_mode, int amplitude_fixed, const int *inverted,
                                      const double *pk_n2) {
  int rc = batch_members(members, count);
  if (rc) return rc;
  if (!seeds || !inverted || !pk_n2) return mgpu_fail(members[0], MGPU_E_INVALID, "null argument");
  for (int r = 0; r < count; r++)
    if (members[r]->fft.tcols != members[0]->fft.tcols)
      return mgpu_fail(members[0], MGPU_E_STATE, "the members' FFT tile widths differ");
  return ic_displacement_fields_batch(members, count, seeds, sphere_mode, amplitude_fixed, inverted, pk_n2);
}

int mgpu_batch_step(mgpu_ctx *const *members, int count, const mgpu_mg *mg, double dda, double A, double s1k,
                    double s2k, double dyyy, double s1d, double s2d, int graph) {
  int rc = batch_members(members, count);
  if (rc) return rc;
  mgpu_ctx *m0 = members[0];
  for (int r = 0; r < count; r++) {
    mgpu_ctx *m = members[r];
    if ((rc = step_check(m, mg))) return mgpu_fail(m0, rc, "member " + std::to_string(r) + ": " + m->err);
    if (m->fft.tcols != m0->fft.tcols) return mgpu_fail(m0, MGPU_E_STATE, "the members' FFT tile widths differ");
  }
  // realization 0's launches with grid z = count reach every member
  m0->bz = (unsigned)count;
  rc = run_step(m0, mg, dda, A, s1k, s2k, dyyy, s1d, s2d, graph);
  m0->bz = 1;
  if (rc) return rc;
  const size_t rec_off = (const char *)m0->dep_rec - (const char *)m0->dep_scratch;
  const size_t off_off = (const char *)m0->dep_off - (const char *)m0->dep_scratch;
  for (int r = 1; r < count; r++) {
    // the host state the step left in realization 0, for each member
    mgpu_ctx *m = members[r];
    m->have_deposit = m0->have_deposit;
    m->have_density = m0->have_density;
    m->have_forces = m0->have_forces;
    m->have_disp = m0->have_disp;
    m->have_phik = m0->have_phik;
    m->dx_slot = m0->dx_slot;
    m->phi_slot = m0->phi_slot;
    for (int a = 0; a < 3; a++) m->force_slot[a] = m0->force_slot[a];
    m->dep_rec = (float4 *)((char *)m->dep_scratch + rec_off);
    m->dep_off = (unsigned *)((char *)m->dep_scratch + off_off);
    m->sorted_valid = m0->sorted_valid;
    m->vsum_valid = m0->vsum_valid;
    m->dep_tiles_grid = m0->dep_tiles_grid;
  }
  return MGPU_OK;
}

int mgpu_get_means(mgpu_ctx *ctx, double sumD[3], double sumV[3]) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  double s[6];
  MGPU_HIP(ctx, hipMemcpyAsync(s, ctx->dsum, sizeof s, hipMemcpyDeviceToHost, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const double n = (double)std::max<size_t>(ctx->np, 1);
  for (int a = 0; a < 3; a++) {
    if (sumD) sumD[a] = s[a] / n;
    if (sumV) sumV[a] = s[3 + a] / n;
  }
  return MGPU_OK;
}

int mgpu_ic_displacement_fields(mgpu_ctx *ctx, unsigned long seed, int sphere_mode, int amplitude_fixed, int inverted,
                                const double *pk_n2) {
  if (check_ctx(ctx) || !pk_n2) return MGPU_E_INVALID;
  return ic_displacement_fields(ctx, seed, sphere_mode, amplitude_fixed, inverted, pk_n2);
}

int mgpu_ic_displacement_fields_noalloc(mgpu_ctx *ctx, unsigned long seed, int sphere_mode, int amplitude_fixed,
                                        int inverted, const double *pk_n2) {
  if (check_ctx(ctx) || !pk_n2) return MGPU_E_INVALID;
  if ((size_t)ctx->N * ctx->N > (size_t)MGPU_IC_SCRATCH_SEEDS)
    return mgpu_fail(ctx, MGPU_E_INVALID, "Nmesh^2 = " + std::to_string((size_t)ctx->N * ctx->N) +
                                              " seeds exceed the scratch of " + std::to_string(MGPU_IC_SCRATCH_SEEDS));
  if (!ctx->ic_scratch) return mgpu_fail(ctx, MGPU_E_STATE, "seed-table scratch not set up");
  return ic_displacement_fields(ctx, seed, sphere_mode, amplitude_fixed, inverted, pk_n2, ctx->ic_scratch);
}

int mgpu_init_particles(mgpu_ctx *ctx, double Di, double Di2) {
  if (ctx) ctx->sorted_valid = false, ctx->vsum_valid = false;
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  int rc = ic_init_particles(ctx, Di, Di2);
  if (rc) return rc;
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  return MGPU_OK;
}

int mgpu_lpt_fields(mgpu_ctx *ctx, int slot, const double *g1_n2, const double *g2_n2) {
  if (check_ctx(ctx) || slot < 0 || slot > 1 || !g1_n2 || !g2_n2) return MGPU_E_INVALID;
  return ic_lpt_fields(ctx, slot, g1_n2, g2_n2);
}

int mgpu_lpt_fields_slices(mgpu_ctx *ctx, int slot, const mgpu_kslice *o1, const mgpu_kslice *o2) {
  if (check_ctx(ctx) || slot < 0 || slot > 1 || !o1 || !o2 || !o1->num || !o1->dnum || !o1->den || !o1->dden ||
      !o2->num || !o2->dnum || !o2->den || !o2->dden)
    return mgpu_fail(ctx, MGPU_E_INVALID, "bad arguments to mgpu_lpt_fields_slices");
  return ic_lpt_fields_slices(ctx, slot, o1, o2);
}

static bool slice_ok(const mgpu_kslice *q) { return q && q->num && q->dnum && q->den && q->dden; }

int mgpu_lpt_fields_slices2(mgpu_ctx *ctx, const mgpu_kslice *s0o1, const mgpu_kslice *s0o2,
                            const mgpu_kslice *s1o1, const mgpu_kslice *s1o2) {
  if (check_ctx(ctx) || !slice_ok(s0o1) || !slice_ok(s0o2) || !slice_ok(s1o1) || !slice_ok(s1o2))
    return mgpu_fail(ctx, MGPU_E_INVALID, "bad arguments to mgpu_lpt_fields_slices2");
  const int slots[2] = {0, 1};
  const mgpu_kslice *q[4] = {s0o1, s0o2, s1o1, s1o2};
  if (!ctx->lattice_slot) {
    int rc = ic_lpt_fields_slices_n(ctx, 1, slots, q);
    return rc ? rc : ic_lpt_fields_slices_n(ctx, 1, slots + 1, q + 2);
  }
  return ic_lpt_fields_slices_n(ctx, 2, slots, q);
}

// ---- mesh inspection: convert between the device layout and the FFTW layout (no ghost slab)
static float2 *mesh_ptr(mgpu_ctx *ctx, int mesh, bool &is_real) {
  is_real = false;
  switch (mesh) {
    case MGPU_MESH_DENSITY_X: is_real = true; return ctx->slot[ctx->dx_slot];
    case MGPU_MESH_DENSITY_K: return ctx->slot[0];
    case MGPU_MESH_PHI_K: return ctx->slot[ctx->phi_slot];
    case MGPU_MESH_FORCE_X: is_real = true; return ctx->slot[ctx->force_slot[0]];
    case MGPU_MESH_FORCE_Y: is_real = true; return ctx->slot[ctx->force_slot[1]];
    case MGPU_MESH_FORCE_Z: is_real = true; return ctx->slot[ctx->force_slot[2]];
    case MGPU_MESH_CDELTA1: return ctx->cdelta[0] ? ctx->cdelta[0] : ctx->slot[0];
    case MGPU_MESH_CDELTA2: return ctx->cdelta[1];
    default: return nullptr;
  }
}

int mgpu_get_mesh(mgpu_ctx *ctx, int mesh, double *host) {
  if (check_ctx(ctx) || !host) return MGPU_E_INVALID;
  bool is_real;
  float2 *src = mesh_ptr(ctx, mesh, is_real);
  if (!src) return mgpu_fail(ctx, MGPU_E_INVALID, "mesh not available");
  std::vector<float2> h(ctx->slot_elems);
  MGPU_HIP(ctx, hipMemcpyAsync(h.data(), src, h.size() * sizeof(float2), hipMemcpyDeviceToHost, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  const size_t N = ctx->N, HP = ctx->HP, H = N / 2 + 1;
  const float *hf = (const float *)h.data();
  for (size_t row = 0; row < N * N; row++) {
    if (is_real)
      for (size_t z = 0; z < N; z++) host[row * N + z] = hf[row * 2 * HP + z];
    else
      for (size_t k = 0; k < H; k++) {
        host[2 * (row * H + k)] = h[row * HP + k].x;
        host[2 * (row * H + k) + 1] = h[row * HP + k].y;
      }
  }
  return MGPU_OK;
}

int mgpu_set_mesh(mgpu_ctx *ctx, int mesh, const double *host) {
  if (check_ctx(ctx) || !host) return MGPU_E_INVALID;
  bool is_real;
  float2 *dst = mesh_ptr(ctx, mesh, is_real);
  if (!dst) return mgpu_fail(ctx, MGPU_E_INVALID, "mesh not available");
  std::vector<float2> h(ctx->slot_elems, make_float2(0.f, 0.f));
  const size_t N = ctx->N, HP = ctx->HP, H = N / 2 + 1;
  float *hf = (float *)h.data();
  for (size_t row = 0; row < N * N; row++) {
    if (is_real)
      for (size_t z = 0; z < N; z++) hf[row * 2 * HP + z] = (float)host[row * N + z];
    else
      for (size_t k = 0; k < H; k++)
        h[row * HP + k] = make_float2((float)host[2 * (row * H + k)], (float)host[2 * (row * H + k) + 1]);
  }
  MGPU_HIP(ctx, hipMemcpyAsync(dst, h.data(), h.size() * sizeof(float2), hipMemcpyHostToDevice, ctx->stream));
  MGPU_HIP(ctx, hipStreamSynchronize(ctx->stream));
  if (mesh == MGPU_MESH_DENSITY_X) ctx->have_deposit = true;  // a density to transform (FFTW's r2c input)
  if (mesh == MGPU_MESH_DENSITY_K) ctx->have_density = true;
  if (mesh == MGPU_MESH_PHI_K) ctx->have_phik = true;
  if (mesh >= MGPU_MESH_FORCE_X && mesh <= MGPU_MESH_FORCE_Z) ctx->have_forces = true;
  return MGPU_OK;
}

int mgpu_get_timers(mgpu_ctx *ctx, mgpu_timers *t, int reset) {
  if (check_ctx(ctx) || !t) return MGPU_E_INVALID;
  *t = ctx->t;
  if (reset) ctx->t = mgpu_timers{};
  return MGPU_OK;
}

int mgpu_enable_timers(mgpu_ctx *ctx, int on) {
  if (check_ctx(ctx)) return MGPU_E_INVALID;
  ctx->timers_on = on != 0;
  return MGPU_OK;
}

}  // extern "C"
