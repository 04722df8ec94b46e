/* A plain C11 program against include/dsm.h: the calls a C host would make (tests/test_host.py
   compiles and links it against libgpu_dsm.so and runs it; without a device it stops at
   dsm_create, which reports the missing device). */
#include <stdio.h>

#include "dsm.h"

int main(void) {
  dsm_engine* e = NULL;
  if (dsm_create(0, &e) != DSM_OK) {
    printf("dsm_create: %s\n", dsm_last_error());
    return 0;
  }
  dsm_ensemble_desc d = {0};
  d.n_chains = 1024;
  d.nk = 50;
  d.beta = 1.0f;
  d.cd = 1;
  d.seed = 1;
  d.sample_mode = DSM_SAMPLE_GT;
  d.delta = 1.0;
  d.corr_p = 16;
  d.corr_m = 2;
  int32_t ens = -1;
  if (dsm_add_ensemble(e, &d, &ens) != DSM_OK || dsm_init_equilibrium(e, ens) != DSM_OK) return 1;
  dsm_run_stats st;
  if (dsm_advance(e, ens, 50.0, DSM_MODE_BATCHED, NULL, NULL, &st) != DSM_OK) return 1;
  double sigma[8];
  int32_t healthy = 0;
  if (dsm_stress(e, ens, sigma, &healthy) != DSM_OK) return 1;
  double lag[64], g[64];
  int32_t np = 0;
  if (dsm_correlator(e, ens, 64, lag, g, &np) != DSM_OK) return 1;
  printf("jumps %llu, healthy %d, <Z> %.3f, G(0) %.4f over %d points\n", (unsigned long long)st.events, healthy,
         sigma[7], g[0], np);
  dsm_destroy(e);
  return 0;
}
