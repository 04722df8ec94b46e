// fake_engine.cpp -- a host-only stand-in for the engine half of the C ABI (include/dsm.h),
// so that the host code above it (the drop-in driver dsm_driver.cpp and the serving
// scheduler dsm_serve.cpp) can be built with g++ and run under AddressSanitizer and
// UndefinedBehaviorSanitizer on a machine without a GPU.  Test infrastructure only: it
// keeps a valid chain state (Z, N summing to Nk, Q with zero ends) and advances clocks,
// but makes no KMC jumps.
#include <cmath>
#include <cstring>
#include <string>
#include <vector>

#include "dsm.h"

struct FakeEns {
  bool live = false;
  dsm_ensemble_desc d{};
  int n = 0, cap = 0, s = 0;
  std::vector<int32_t> Z, N, stall;
  std::vector<float> Q, w, dt;
  std::vector<double> t;
  std::vector<uint64_t> ev;
  std::vector<int64_t> ns;
  std::vector<double> rows;
  int64_t k_cur = 0;
};
struct dsm_engine {
  std::vector<FakeEns> ens;
};

static thread_local std::string g_fake_err;
static int fake_fail(int rc, const char* m) {
  g_fake_err = m;
  return rc;
}

extern "C" {
const char* dsm_last_error(void) { return g_fake_err.c_str(); }
int dsm_set_option(dsm_engine* e, int32_t, int32_t) { return e ? DSM_OK : DSM_ERR_ARG; }
int dsm_mem_info(int32_t, size_t* f, size_t* t) {
  if (f) *f = (size_t)64 << 30;
  if (t) *t = (size_t)288 << 30;
  return DSM_OK;
}
int dsm_create(int32_t device_id, dsm_engine** out) {
  if (device_id != 0) return fake_fail(DSM_ERR_DEVICE, "no such device");
  *out = new dsm_engine();
  return DSM_OK;
}
int dsm_destroy(dsm_engine* e) {
  delete e;
  return DSM_OK;
}
int dsm_add_ensemble(dsm_engine* e, const dsm_ensemble_desc* d, int32_t* id) {
  if (!e || !d || d->n_chains <= 0 || d->nk < 2) return fake_fail(DSM_ERR_ARG, "bad ensemble");
  FakeEns E;
  E.live = true;
  E.d = *d;
  E.n = d->n_chains;
  E.cap = d->nk;
  E.s = d->corr_levels > 0 ? d->corr_levels : 1;
  E.Z.assign(E.n, 1);
  E.N.assign((size_t)E.n * E.cap, 0);
  E.Q.assign((size_t)E.n * E.cap * 3, 0.0f);
  E.w.assign((size_t)E.n * E.cap, 0.0f);
  E.stall.assign(E.n, 0);
  E.dt.assign(E.n, 0.0f);
  E.t.assign(E.n, 0.0);
  E.ev.assign(E.n, 0);
  E.ns.assign(E.n, 0);
  E.k_cur = d->sample_mode == DSM_SAMPLE_FLOW ? 1 : 0;
  e->ens.push_back(E);
  *id = (int32_t)e->ens.size() - 1;
  return DSM_OK;
}
static FakeEns* get(dsm_engine* e, int id) {
  return (e && id >= 0 && id < (int)e->ens.size() && e->ens[id].live) ? &e->ens[id] : nullptr;
}
int dsm_remove_ensemble(dsm_engine* e, int32_t id) {
  FakeEns* E = get(e, id);
  if (!E) return DSM_ERR_ARG;
  E->live = false;
  E->Z.clear(); E->N.clear(); E->Q.clear(); E->w.clear();
  return DSM_OK;
}
int dsm_init_equilibrium(dsm_engine* e, int32_t id) {
  FakeEns* E = get(e, id);
  if (!E) return DSM_ERR_ARG;
  const int nk = E->d.nk;
  for (int i = 0; i < E->n; ++i) {  // three strands: 1, nk-2, 1 Kuhn steps (ends Q = 0)
    const int Z = nk >= 3 ? 3 : 1;
    E->Z[i] = Z;
    int32_t* N = &E->N[(size_t)i * E->cap];
    float* Q = &E->Q[(size_t)i * E->cap * 3];
    if (Z == 3) {
      N[0] = 1; N[1] = nk - 2; N[2] = 1;
      Q[3] = 0.1f * (float)(i % 7); Q[4] = -0.2f; Q[5] = 0.3f;
      E->w[(size_t)i * E->cap] = 0.01f;
      E->w[(size_t)i * E->cap + 1] = 0.02f;
    } else {
      N[0] = nk;
    }
  }
  return DSM_OK;
}
int dsm_state_cap(dsm_engine* e, int32_t id, int32_t* cap) {
  FakeEns* E = get(e, id);
  if (!E || !cap) return DSM_ERR_ARG;
  *cap = E->cap;
  return DSM_OK;
}
int dsm_ensemble_info(dsm_engine* e, int32_t id, int32_t* n, int32_t* cap, int32_t* mode) {
  FakeEns* E = get(e, id);
  if (!E) return DSM_ERR_ARG;
  if (n) *n = E->n;
  if (cap) *cap = E->cap;
  if (mode) *mode = E->d.sample_mode;
  return DSM_OK;
}
int dsm_get_state(dsm_engine* e, int32_t id, int32_t cap, int32_t* Z, int32_t* N, float* Q, float* w, double* t,
                  uint64_t* ev, int32_t* stall, float* dt_prev, int64_t* ns) {
  FakeEns* E = get(e, id);
  if (!E || cap < E->cap) return DSM_ERR_ARG;
  for (int i = 0; i < E->n; ++i) {
    if (Z) Z[i] = E->Z[i];
    if (t) t[i] = E->t[i];
    if (ev) ev[i] = E->ev[i];
    if (stall) stall[i] = E->stall[i];
    if (dt_prev) dt_prev[i] = E->dt[i];
    if (ns) ns[i] = E->ns[i];
    for (int k = 0; k < cap; ++k) {
      const bool in = k < E->cap;
      const size_t o = (size_t)i * cap + k, s = (size_t)i * E->cap + k;
      if (N) N[o] = in ? E->N[s] : 0;
      if (w) w[o] = in ? E->w[s] : 0.0f;
      for (int c = 0; c < 3 && Q; ++c) Q[o * 3 + c] = in ? E->Q[s * 3 + c] : 0.0f;
    }
  }
  return DSM_OK;
}
int dsm_set_state(dsm_engine* e, int32_t id, int32_t cap, const int32_t* Z, const int32_t* N, const float* Q,
                  const float* w, const double* t, const uint64_t* ev, const int32_t* stall, const float* dt_prev,
                  const int64_t* ns) {
  FakeEns* E = get(e, id);
  if (!E || !Z || !N || !Q || !w || cap < E->cap) return fake_fail(DSM_ERR_ARG, "bad state");
  for (int i = 0; i < E->n; ++i) {
    if (Z[i] < 1 || Z[i] > E->d.nk) return fake_fail(DSM_ERR_ARG, "Z out of range");
    int64_t sum = 0;
    for (int k = 0; k < Z[i]; ++k) {
      if (N[(size_t)i * cap + k] < 1) return fake_fail(DSM_ERR_ARG, "N < 1");
      sum += N[(size_t)i * cap + k];
    }
    if (sum != E->d.nk) return fake_fail(DSM_ERR_ARG, "sum N != Nk");
  }
  for (int i = 0; i < E->n; ++i) {
    E->Z[i] = Z[i];
    E->t[i] = t ? t[i] : 0.0;
    E->ev[i] = ev ? ev[i] : 0;
    E->stall[i] = stall ? stall[i] : 0;
    E->dt[i] = dt_prev ? dt_prev[i] : 0.0f;
    E->ns[i] = ns ? ns[i] : 0;
    for (int k = 0; k < E->cap; ++k) {
      const size_t o = (size_t)i * cap + k, s = (size_t)i * E->cap + k;
      E->N[s] = k < Z[i] ? N[o] : 0;
      E->w[s] = k < Z[i] ? w[o] : 0.0f;
      for (int c = 0; c < 3; ++c) E->Q[s * 3 + c] = k < Z[i] ? Q[o * 3 + c] : 0.0f;
    }
  }
  if (ns) E->k_cur = ns[0];
  return DSM_OK;
}
static int fake_advance(FakeEns* E, double t_end) {
  for (int i = 0; i < E->n; ++i)
    if (E->t[i] < t_end) { E->ev[i] += (uint64_t)std::ceil(t_end - E->t[i]); E->t[i] = t_end; }
  if (E->d.sample_mode == DSM_SAMPLE_FLOW) {
    const int64_t k_target = (int64_t)std::ceil(t_end / E->d.delta - 1e-12) + 1;
    for (; E->k_cur < k_target; ++E->k_cur)
      for (int q = 0; q < 9; ++q) E->rows.push_back(q == 8 ? (double)E->n : -1.0 * E->n);
  }
  return DSM_OK;
}
int dsm_advance(dsm_engine* e, int32_t id, double t_end, int32_t, volatile const int32_t* run_flag,
                volatile int32_t* progress, dsm_run_stats* st) {
  FakeEns* E = get(e, id);
  if (!E) return DSM_ERR_ARG;
  if (run_flag && *run_flag == 0) return DSM_ABORTED;
  fake_advance(E, t_end);
  if (progress) *progress = 100;
  if (st) std::memset(st, 0, sizeof(*st));
  return DSM_OK;
}
int dsm_advance_many(dsm_engine* e, const int32_t* ids, const double* t_end, int32_t count, int32_t,
                     volatile const int32_t*, dsm_run_stats* st) {
  if (st) std::memset(st, 0, sizeof(*st));
  for (int q = 0; q < count; ++q) {
    FakeEns* E = get(e, ids[q]);
    if (!E) return DSM_ERR_ARG;
    fake_advance(E, t_end[q]);
    if (st) st->events += (uint64_t)E->n;
  }
  return DSM_OK;
}
int dsm_flow_stress(dsm_engine* e, int32_t id, int64_t first, int64_t max_rows, double* out, int64_t* n_rows) {
  FakeEns* E = get(e, id);
  if (!E) return DSM_ERR_ARG;
  const int64_t have = (int64_t)E->rows.size() / 9;
  if (n_rows) *n_rows = have;
  for (int64_t r = first; out && r < have && r - first < max_rows; ++r)
    std::memcpy(out + (r - first) * 9, &E->rows[(size_t)r * 9], 9 * sizeof(double));
  return DSM_OK;
}
int dsm_correlator(dsm_engine* e, int32_t id, int32_t max_pts, double* lag, double* g, int32_t* npts) {
  FakeEns* E = get(e, id);
  if (!E) return DSM_ERR_ARG;
  const int p = E->d.corr_p > 0 ? E->d.corr_p : 64, m = E->d.corr_m > 0 ? E->d.corr_m : 8;
  const int np = p + (E->s - 1) * (p - p / m);
  for (int i = 0; i < np && i < max_pts; ++i) {
    if (lag) lag[i] = i;
    if (g) g[i] = 1.0 / (1.0 + i);
  }
  if (npts) *npts = np;
  return DSM_OK;
}
}  // extern "C"
