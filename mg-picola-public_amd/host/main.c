/* mgpicola: command-line entry of the MI355X MG-PICOLA path.
 *
 *   mgpicola [--model FOFR|FOFR_LCDM|DGP|BRANSDICKE|MBETA] [--compute-pofk] paramfile
 *
 * mirrors the reference's `MG_PICOLA_<MODEL> paramfile` (src/main.c:50-68):
 * the model and the COMPUTE_POFK switch, compile-time in the reference, are
 * chosen here by flags or by an executable name ending in _<MODEL>. */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "mgpicola.h"

int main(int argc, char **argv) {
  int model = MGP_MODEL_DGP, compute_pofk = 0;
  const char *pf = NULL;
  const char *us = strrchr(argv[0], '_');
  if (us && mgp_model_from_name(us + 1) >= 0) model = mgp_model_from_name(us + 1);
  for (int i = 1; i < argc; i++) {
    if (!strcmp(argv[i], "--model") && i + 1 < argc) {
      model = mgp_model_from_name(argv[++i]);
      if (model < 0) {
        fprintf(stderr, "unknown model '%s'\n", argv[i]);
        return 1;
      }
    } else if (!strcmp(argv[i], "--compute-pofk")) {
      compute_pofk = 1;
    } else {
      pf = argv[i];
    }
  }
  if (!pf) {
    fprintf(stdout, "Input parameters not found\nCall with <ParameterFile>\n");
    return 1;
  }
  mgp_params p;
  mgp_params_defaults(&p, model);
  p.compute_pofk = compute_pofk;
  char err[4096];
  if (mgp_read_paramfile(pf, &p, err, sizeof err) || mgp_read_outputs(&p, err, sizeof err)) {
    fprintf(stdout, "\nERROR: %s\n\n", err);
    return 1;
  }
  mgp_run_stats st;
  if (mgp_run(&p, 1, &st, err, sizeof err)) {
    fprintf(stdout, "\nERROR: %s\n\n", err);
    return 1;
  }
  double ps = (double)p.Nsample * p.Nsample * p.Nsample * st.steps / (st.step_seconds > 0 ? st.step_seconds : 1);
  printf("[mgpicola] %ld steps in %.3f s: %.4g particle-steps/s\n", st.steps, st.step_seconds, ps);
  return 0;
}
