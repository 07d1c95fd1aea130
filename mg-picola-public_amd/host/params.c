/* Paramfile and output-list readers with the reference's strictness
 * (src/read_param.c:32-445, src/user_defined_functions.h:227-406): the tag
 * table depends on the build configuration, every tag must appear exactly
 * once, unknown or repeated tags are errors, '%' starts a comment. */
#include "internal.h"

#include <stdio.h>
#include <stdlib.h>
#include <string.h>

enum { T_FLOAT = 1, T_STRING = 2, T_INT = 3 };
typedef struct {
  const char *name;
  int type;
  void *addr;
} tagdef;

#define MAXTAGS 128

static int build_tags(mgp_params *p, tagdef *t) {
  int n = 0;
#define ADD(nm, ty, ad) t[n++] = (tagdef){nm, ty, (void *)(ad)}
  /* read_mg_parameters */
  ADD("modified_gravity_active", T_INT, &p->modified_gravity_active);
  ADD("use_lcdm_growth_factors", T_INT, &p->use_lcdm_growth_factors);
  ADD("input_pofk_is_for_lcdm", T_INT, &p->input_pofk_is_for_lcdm);
  ADD("input_sigma8_is_for_lcdm", T_INT, &p->input_sigma8_is_for_lcdm);
  ADD("inverted_initial_condition", T_INT, &p->inverted_initial_condition);
  ADD("amplitude_fixed_initial_condition", T_INT, &p->amplitude_fixed_initial_condition);
  ADD("include_screening", T_INT, &p->include_screening);
  if (p->compute_pofk) {
    ADD("pofk_compute_every_step", T_INT, &p->pofk_compute_every_step);
    ADD("pofk_compute_rsd_pofk", T_INT, &p->pofk_compute_rsd_pofk);
    ADD("pofk_nbins", T_INT, &p->pofk_nbins);
    ADD("pofk_bintype", T_INT, &p->pofk_bintype);
    ADD("pofk_subtract_shotnoise", T_INT, &p->pofk_subtract_shotnoise);
    ADD("pofk_kmin", T_FLOAT, &p->pofk_kmin);
    ADD("pofk_kmax", T_FLOAT, &p->pofk_kmax);
  }
  switch (p->model) {
    case MGP_MODEL_FOFRNU:
      ADD("nu_include_massive_neutrinos", T_INT, &p->nu_include_massive_neutrinos);
      ADD("nu_SumMassNuEV", T_FLOAT, &p->nu_SumMassNuEV);
      ADD("nu_FilenameTransferInfofile", T_STRING, p->nu_FilenameTransferInfofile);
      /* and the f(R) tags */
      __attribute__((fallthrough));
    case MGP_MODEL_FOFR:
    case MGP_MODEL_FOFR_LCDM:
      ADD("fofr0", T_FLOAT, &p->fofr0);
      ADD("nfofr", T_FLOAT, &p->nfofr);
      break;
    case MGP_MODEL_MBETA:
      ADD("range_symm", T_FLOAT, &p->range_symm);
      ADD("assb_symm", T_FLOAT, &p->assb_symm);
      ADD("beta_symm", T_FLOAT, &p->beta_symm);
      break;
    case MGP_MODEL_DGP:
      ADD("Rsmooth", T_FLOAT, &p->Rsmooth);
      ADD("rcH0_DGP", T_FLOAT, &p->rcH0_DGP);
      break;
    case MGP_MODEL_BRANSDICKE:
      ADD("wBD", T_FLOAT, &p->wBD);
      ADD("Omegah2", T_FLOAT, &p->Omegah2);
      ADD("Omegar2", T_FLOAT, &p->Omegarh2);
      ADD("Omegav2", T_FLOAT, &p->Omegavh2);
      break;
  }
  /* core parameters */
  ADD("OutputDir", T_STRING, p->OutputDir);
  ADD("FileBase", T_STRING, p->FileBase);
  ADD("OutputRedshiftFile", T_STRING, p->OutputRedshiftFile);
  ADD("NumFilesWrittenInParallel", T_INT, &p->NumFilesWrittenInParallel);
  ADD("UseCOLA", T_INT, &p->UseCOLA);
  ADD("Buffer", T_FLOAT, &p->Buffer);
  ADD("Nmesh", T_INT, &p->Nmesh);
  ADD("Nsample", T_INT, &p->Nsample);
  ADD("Box", T_FLOAT, &p->Box);
  ADD("Init_Redshift", T_FLOAT, &p->Init_Redshift);
  ADD("Seed", T_INT, &p->Seed);
  ADD("SphereMode", T_INT, &p->SphereMode);
  ADD("WhichSpectrum", T_INT, &p->WhichSpectrum);
  ADD("WhichTransfer", T_INT, &p->WhichTransfer);
  ADD("FileWithInputSpectrum", T_STRING, p->FileWithInputSpectrum);
  ADD("FileWithInputTransfer", T_STRING, p->FileWithInputTransfer);
  ADD("Omega", T_FLOAT, &p->Omega);
  ADD("OmegaBaryon", T_FLOAT, &p->OmegaBaryon);
  ADD("HubbleParam", T_FLOAT, &p->HubbleParam);
  ADD("Sigma8", T_FLOAT, &p->Sigma8);
  ADD("PrimordialIndex", T_FLOAT, &p->PrimordialIndex);
  ADD("UnitVelocity_in_cm_per_s", T_FLOAT, &p->UnitVelocity_in_cm_per_s);
  ADD("UnitLength_in_cm", T_FLOAT, &p->UnitLength_in_cm);
  ADD("UnitMass_in_g", T_FLOAT, &p->UnitMass_in_g);
  ADD("InputSpectrum_UnitLength_in_cm", T_FLOAT, &p->InputSpectrum_UnitLength_in_cm);
#undef ADD
  return n;
}

int mgp_model_from_name(const char *name) {
  if (!strcmp(name, "FOFR")) return MGP_MODEL_FOFR;
  if (!strcmp(name, "FOFR_LCDM")) return MGP_MODEL_FOFR_LCDM;
  if (!strcmp(name, "DGP")) return MGP_MODEL_DGP;
  if (!strcmp(name, "BRANSDICKE")) return MGP_MODEL_BRANSDICKE;
  if (!strcmp(name, "MBETA")) return MGP_MODEL_MBETA;
  if (!strcmp(name, "FOFRNU")) return MGP_MODEL_FOFRNU;
  return -1;
}

void mgp_params_defaults(mgp_params *p, int model) {
  memset(p, 0, sizeof(*p));
  p->model = model;
  p->scale_dependent = model == MGP_MODEL_FOFR || model == MGP_MODEL_MBETA || model == MGP_MODEL_FOFRNU;
  p->gadget_style = 1;
  p->growth_on_gpu = 1;
  p->fuse_kick_drift = 1;
  p->write_particles = 1;
}

int mgp_param_tag_count(const mgp_params *p) {
  tagdef t[MAXTAGS];
  return build_tags((mgp_params *)p, t);
}
const char *mgp_param_tag_name(const mgp_params *p, int i) {
  tagdef t[MAXTAGS];
  int n = build_tags((mgp_params *)p, t);
  return (i >= 0 && i < n) ? t[i].name : NULL;
}

/* read_parameterfile, src/read_param.c:107-445 */
int mgp_read_paramfile(const char *fname, mgp_params *p, char *err, size_t errlen) {
  tagdef t[MAXTAGS];
  int seen[MAXTAGS] = {0};
  int nt = build_tags(p, t);
  FILE *fd = fopen(fname, "r");
  if (!fd) {
    snprintf(err, errlen, "Parameter file '%s' not found", fname);
    return -1;
  }
  char buf[500], b1[500], b2[500], b3[500];
  int bad = 0;
  size_t used = 0;
  err[0] = 0;
  while (fgets(buf, sizeof buf, fd)) {
    if (sscanf(buf, "%499s%499s%499s", b1, b2, b3) < 2) continue;
    if (b1[0] == '%') continue;
    int j = -1;
    for (int i = 0; i < nt; i++)
      if (!seen[i] && !strcmp(b1, t[i].name)) {
        j = i;
        break;
      }
    if (j < 0) {
      used += snprintf(err + used, used < errlen ? errlen - used : 0,
                       "Tag '%s' not allowed or multiple defined. ", b1);
      bad = 1;
      continue;
    }
    seen[j] = 1;
    switch (t[j].type) {
      case T_FLOAT: *(double *)t[j].addr = atof(b2); break;
      case T_STRING: strcpy((char *)t[j].addr, b2); break;
      case T_INT: *(int *)t[j].addr = atoi(b2); break;
    }
  }
  fclose(fd);
  for (int i = 0; i < nt; i++)
    if (!seen[i]) {
      used += snprintf(err + used, used < errlen ? errlen - used : 0, "I miss a value for tag '%s'. ", t[i].name);
      bad = 1;
    }
  if (bad) return -1;
  if (p->NumFilesWrittenInParallel < 1) {
    snprintf(err, errlen, "NumFilesWrittenInParallel is less than 1");
    return -1;
  }
  if (p->WhichSpectrum != 0 && p->WhichTransfer != 0) {
    snprintf(err, errlen, "can't run with both power spectrum and transfer function");
    return -1;
  }
  if (p->WhichSpectrum == 0 && p->WhichTransfer == 0) {
    snprintf(err, errlen, "can't run with neither power spectrum nor transfer function");
    return -1;
  }
  if (!p->compute_pofk) {
    /* compute_power_spectrum's fiducial binning without COMPUTE_POFK (src/compute_pofk.c:95-101) */
    p->pofk_nbins = p->Nmesh / 2;
    p->pofk_bintype = 0;
    p->pofk_subtract_shotnoise = 1;
    p->pofk_kmin = 0.0;
    p->pofk_kmax = 2.0 * MGP_PI / p->Box * p->Nmesh / 2.0;
  }
  return 0;
}

static int cmp_out(const void *a, const void *b) {
  double x = *(const double *)a, y = *(const double *)b;
  return x > y ? -1 : x < y ? 1 : 0;
}

/* read_outputs, src/read_param.c:32-100 */
int mgp_read_outputs(mgp_params *p, char *err, size_t errlen) {
  FILE *fd = fopen(p->OutputRedshiftFile, "r");
  if (!fd) {
    snprintf(err, errlen, "Output Redshift File '%s' not found", p->OutputRedshiftFile);
    return -1;
  }
  char buf[500];
  double pairs[2 * MGP_MAXOUT];
  int n = 0;
  while (fgets(buf, sizeof buf, fd)) {
    int nsteps;
    double red;
    if (buf[0] == '%') continue;
    int got = sscanf(buf, "%lf, %d", &red, &nsteps);
    if (got > 0) {
      if (got != 2) {
        fclose(fd);
        snprintf(err, errlen, "Line in Output Redshift File '%s' is in incorrect format", buf);
        return -1;
      }
      if (nsteps <= 0 && (p->Init_Redshift - red) / p->Init_Redshift > 1.0E-6) {
        fclose(fd);
        snprintf(err, errlen, "read a value for nsteps of <= 0 up to redshift %lf", red);
        return -1;
      }
      if (n == MGP_MAXOUT) {
        fclose(fd);
        snprintf(err, errlen, "too many output redshifts");
        return -1;
      }
      pairs[2 * n] = red;
      pairs[2 * n + 1] = nsteps;
      n++;
    }
  }
  fclose(fd);
  if (n == 0) {
    snprintf(err, errlen, "Found no output redshifts in file '%s'", p->OutputRedshiftFile);
    return -1;
  }
  qsort(pairs, n, 2 * sizeof(double), cmp_out);
  for (int i = 1; i < n; i++)
    if (pairs[2 * i] == pairs[2 * (i - 1)]) {
      snprintf(err, errlen, "Duplicate output redshift (%lf) in Output Redshift File", pairs[2 * i]);
      return -1;
    }
  if (pairs[0] > p->Init_Redshift) {
    snprintf(err, errlen, "The highest output redshift (%lf) is greater than the initial redshift (%lf)", pairs[0],
             p->Init_Redshift);
    return -1;
  }
  p->Noutputs = n;
  for (int i = 0; i < n; i++) {
    p->out_z[i] = pairs[2 * i];
    p->out_nsteps[i] = (int)pairs[2 * i + 1];
  }
  return 0;
}
