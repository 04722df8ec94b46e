/* oracle_checks.c -- drives the CPU oracle through its threaded paths (ensembles advanced
   on several threads with every sampling mode, partner backgrounds, the long-horizon G(t)
   runner, the timing runner) for runs under AddressSanitizer + UndefinedBehaviorSanitizer
   and under ThreadSanitizer.  Exit code 0 when every run keeps the model's invariants. */
#include <stdio.h>
#include <stdlib.h>

#include "dsm_oracle.h"

int main(void) {
  int bad = 0;
  const float kappa[9] = {0.01f, 0.05f, 0, -0.02f, 0, 0.03f, 0, 0, -0.01f};
  const float zero[9] = {0};
  const float gex[4] = {1.5f, 1.0f, 40000.0f, 120.0f};
  int nk_mix[64];
  for (int i = 0; i < 64; ++i) nk_mix[i] = 12 + (i * 37) % 300;
  for (int mode = 0; mode < 5; ++mode) {
    const int nk = mode == 4 ? 320 : 40;
    orc_params* p = orc_params_new(1.0f, nk, nk, 1, mode == 1 ? kappa : zero, mode == 3);
    if (mode == 2) orc_params_set_msd(p, 1);
    if (mode == 4 && orc_params_set_partner(p, 2, gex, nk_mix, 64)) bad++;
    if (mode == 3 && orc_params_set_partner(p, 1, gex, NULL, 0)) bad++;
    for (int fp64 = 0; fp64 < 2; ++fp64) {
      const int n = 64;
      orc_ens* e = orc_ens_new(p, 7, 1, 0, n, mode == 4 ? nk_mix : NULL, fp64);
      const int ncomp = mode == 1 ? 8 : 3;
      orc_params_set_flow_samples(p, mode == 1);
      void* smp = malloc((size_t)20 * n * ncomp * (fp64 ? sizeof(double) : sizeof(float)));
      orc_ens_advance(e, 20.0, 1 << 30, 1.0, mode == 1 ? 1 : 0, 20, smp, 4);
      bad += orc_ens_invariants(e);
      double* st = malloc(sizeof(double) * 8 * n);
      orc_ens_stress(e, st);
      free(st);
      free(smp);
      orc_ens_delete(e);
    }
    orc_params_delete(p);
  }
  orc_params* p = orc_params_new(1.0f, 25, 25, 0, zero, 0);
  double* g = malloc(sizeof(double) * 32 * 200);
  double sec = 0.0;
  const uint32_t key[2] = {1, 0};
  if (orc_gt_chains(p, key, 32, 0, NULL, 1.0, 2000, 64, 16, 2, 8, 4, 1, g, &sec) <= 0) bad++;
  if (orc_bench(p, key, 64, 0, 50.0, 4, 0, &sec) <= 0) bad++;
  free(g);
  orc_params_delete(p);
  printf("oracle checks: %d failures\n", bad);
  return bad ? 1 : 0;
}
