// dsm_serve.cpp -- serving: a stream of simulation requests on one engine.
//
// The reference runs one job per process (common/main_cuda.cu:34-185, one input.dat);
// several jobs share a GPU only as separate processes.  Here a seeded synthetic
// request stream (Poisson arrivals; Nk, beta, chain count and shear rate drawn per
// request) is admitted into one engine while the requests' estimated device memory
// fits a share of the free HBM, all admitted requests are advanced together, one
// scheduling window at a time, with dsm_advance_many (one KMC pass per window over
// every active ensemble), and each request's latency is measured from its arrival to
// the retrieval of its result (G(t) or the flow stress rows).
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <string>
#include <thread>
#include <vector>

#include "dsm.h"

namespace {

struct Req {
  double arrival = 0.0;  // s after the start
  int nk = 0;
  float beta = 1.0f;
  int nch = 0;
  double rate = 0.0;     // shear rate (0: equilibrium G(t))
  int64_t bytes = 0;
  int ens = -1;
  double t = 0.0;        // tau_K reached
  double done = -1.0;    // s after the start
};

uint64_t splitmix64(uint64_t& x) {
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double uniform(uint64_t& s) { return (double)(splitmix64(s) >> 11) * 0x1p-53; }

bool valid(const dsm_serve_config* c) {
  return c && c->n_requests > 0 && c->arrival_rate > 0.0 && c->nk_min >= 2 && c->nk_max >= c->nk_min &&
         c->nk_max <= 2047 && c->nk_step > 0 && c->n_beta >= 1 && c->n_beta <= 4 && c->nch_min >= 1 &&
         c->nch_max >= c->nch_min && c->flow_fraction >= 0.0 && c->flow_fraction <= 1.0 &&
         (c->flow_fraction == 0.0 || (c->rate_min > 0.0 && c->rate_max >= c->rate_min)) && c->horizon > 0.0 &&
         c->window > 0.0 && c->hbm_fraction > 0.0 && c->hbm_fraction <= 1.0 && c->progress_s >= 0.0;
}

constexpr int kCorrP = 16, kCorrM = 2, kStageRows = 16;

int corr_levels(double horizon) {
  int s = (int)std::ceil(std::log(horizon / kCorrP) / std::log((double)kCorrM)) + 1;
  return s < 1 ? 1 : s;
}

// device bytes of one request: strand records, per-chain header, sample staging and,
// for G(t), the FP64 correlator registers
int64_t estimate_bytes(const Req& r, double horizon) {
  const int64_t n = r.nch, cap = (r.nk + 3) & ~3;
  int64_t b = n * cap * 20 + n * 96;
  if (r.rate > 0.0) {
    b += (int64_t)kStageRows * n * 8 * 4;
  } else {
    const int64_t s = corr_levels(horizon);
    b += (int64_t)kStageRows * n * 3 * 4 + n * 8 * (s * kCorrP * 3 + s * 3 + s * kCorrP);
  }
  return b;
}

std::vector<Req> plan(const dsm_serve_config* c) {
  std::vector<Req> out((size_t)c->n_requests);
  uint64_t st = c->seed ^ 0x5DEECE66Dull;
  double t = 0.0;
  const int nsteps = (c->nk_max - c->nk_min) / c->nk_step + 1;
  for (auto& r : out) {
    t += -std::log(1.0 - uniform(st)) / c->arrival_rate;
    r.arrival = t;
    r.nk = c->nk_min + c->nk_step * std::min(nsteps - 1, (int)(uniform(st) * nsteps));
    r.beta = c->beta[std::min(c->n_beta - 1, (int)(uniform(st) * c->n_beta))];
    const double lo = std::log((double)c->nch_min), hi = std::log((double)c->nch_max + 1.0);
    r.nch = std::min(c->nch_max, (int)std::exp(lo + (hi - lo) * uniform(st)));
    const bool flow = uniform(st) < c->flow_fraction;
    const double u = uniform(st);
    r.rate = flow ? std::exp(std::log(c->rate_min) + (std::log(c->rate_max) - std::log(c->rate_min)) * u) : 0.0;
    r.bytes = estimate_bytes(r, c->horizon);
  }
  return out;
}

double quantile(std::vector<double> v, double q) {
  if (v.empty()) return 0.0;
  std::sort(v.begin(), v.end());
  const double pos = q * (double)(v.size() - 1);
  const size_t i = (size_t)pos;
  const double f = pos - (double)i;
  return i + 1 < v.size() ? v[i] * (1.0 - f) + v[i + 1] * f : v[i];
}

}  // namespace

extern "C" int dsm_serve_plan(const dsm_serve_config* cfg, int32_t max_requests, double* arrival_s, int32_t* nk,
                              float* beta, int32_t* n_chains, double* shear_rate, int64_t* bytes) {
  if (!valid(cfg)) return DSM_ERR_ARG;
  const std::vector<Req> rq = plan(cfg);
  for (int i = 0; i < (int)rq.size() && i < max_requests; ++i) {
    if (arrival_s) arrival_s[i] = rq[i].arrival;
    if (nk) nk[i] = rq[i].nk;
    if (beta) beta[i] = rq[i].beta;
    if (n_chains) n_chains[i] = rq[i].nch;
    if (shear_rate) shear_rate[i] = rq[i].rate;
    if (bytes) bytes[i] = rq[i].bytes;
  }
  return DSM_OK;
}

extern "C" int dsm_serve(int32_t device_id, const dsm_serve_config* cfg, dsm_serve_report* report,
                         double* latency_s) {
  return dsm_serve_ex(device_id, cfg, report, latency_s, nullptr);
}

extern "C" int dsm_serve_ex(int32_t device_id, const dsm_serve_config* cfg, dsm_serve_report* report,
                            double* latency_s, double* summary) {
  if (!valid(cfg) || !report) return DSM_ERR_ARG;
  std::vector<Req> rq = plan(cfg);
  dsm_engine* e = nullptr;
  int rc = dsm_create(device_id, &e);
  if (rc) return rc;
  if (cfg->graph && (rc = dsm_set_option(e, DSM_OPT_GRAPH, 1))) { dsm_destroy(e); return rc; }
  size_t free_b = 0, total_b = 0;
  if ((rc = dsm_mem_info(device_id, &free_b, &total_b))) { dsm_destroy(e); return rc; }
  const int64_t budget = (int64_t)((double)free_b * cfg->hbm_fraction);
  std::memset(report, 0, sizeof(*report));
  using clk = std::chrono::steady_clock;
  const auto t0 = clk::now();
  auto now = [&]() { return std::chrono::duration<double>(clk::now() - t0).count(); };
  size_t next = 0;           // first request not yet admitted (admission in arrival order)
  std::vector<int> active;   // indices into rq
  int64_t used = 0;
  int completed = 0;
  std::vector<double> corr_lag, corr_g, flow_rows;
  double next_note = cfg->progress_s;
  auto note = [&](double tn) {
    if (cfg->progress_s <= 0.0 || tn < next_note) return;
    const size_t arrived = (size_t)(std::upper_bound(rq.begin(), rq.end(), tn,
                                                     [](double x, const Req& r) { return x < r.arrival; }) -
                                    rq.begin());
    std::fprintf(stderr, "serve: %.1f s  arrived %zu  admitted %zu  completed %d  active %zu  rounds %d  jumps %.3g\n",
                 tn, arrived, next, completed, active.size(), report->rounds, (double)report->jumps);
    std::fflush(stderr);
    next_note = tn + cfg->progress_s;
  };
  while (completed < (int)rq.size()) {
    const double tn = now();
    note(tn);
    // admission: arrived requests, in order, while their memory fits
    while (next < rq.size() && rq[next].arrival <= tn) {
      Req& r = rq[next];
      if (r.bytes > budget) { rc = DSM_ERR_MEMORY; break; }
      if (used + r.bytes > budget) break;
      dsm_ensemble_desc d{};
      d.n_chains = r.nch;
      d.nk = r.nk;
      d.beta = r.beta;
      d.cd = cfg->cd;
      d.kappa[1] = (float)r.rate;
      d.seed = (uint32_t)cfg->seed;
      d.salt = (uint32_t)next;
      d.delta = 1.0;
      d.window = kStageRows;
      if (r.rate > 0.0) {
        d.sample_mode = DSM_SAMPLE_FLOW;
      } else {
        d.sample_mode = DSM_SAMPLE_GT;
        d.corr_p = kCorrP;
        d.corr_m = kCorrM;
        d.corr_levels = corr_levels(cfg->horizon);
      }
      const double ta = now();
      if ((rc = dsm_add_ensemble(e, &d, &r.ens)) || (rc = dsm_init_equilibrium(e, r.ens))) break;
      report->admit_s += now() - ta;
      note(now());
      used += r.bytes;
      report->peak_bytes = std::max<int64_t>(report->peak_bytes, used);
      active.push_back((int)next);
      ++next;
    }
    if (rc) break;
    report->max_active = std::max<int32_t>(report->max_active, (int32_t)active.size());
    if (active.empty()) {
      if (next < rq.size()) {
        const double wait = rq[next].arrival - now();
        const double nap = cfg->progress_s > 0.0 ? std::min(wait, cfg->progress_s) : wait;
        if (nap > 0) std::this_thread::sleep_for(std::chrono::duration<double>(nap));
      }
      continue;
    }
    // one scheduling round: every active request advances by one window
    std::vector<int32_t> ids;
    std::vector<double> te;
    for (int i : active) {
      ids.push_back(rq[i].ens);
      te.push_back(std::min(cfg->horizon, rq[i].t + cfg->window));
    }
    dsm_run_stats st{};
    const double tr = now();
    if ((rc = dsm_advance_many(e, ids.data(), te.data(), (int32_t)ids.size(), DSM_MODE_BATCHED, nullptr, &st)))
      break;
    const double tf = now();
    report->advance_s += tf - tr;
    report->jumps += st.events;
    report->rounds++;
    std::vector<int> still;
    for (size_t q = 0; q < active.size(); ++q) {
      Req& r = rq[active[q]];
      r.t = te[q];
      if (r.t < cfg->horizon) { still.push_back(active[q]); continue; }
      // result retrieval is part of the request
      double* sm = summary ? summary + (size_t)active[q] * DSM_SERVE_SUMMARY : nullptr;
      if (sm) std::fill(sm, sm + DSM_SERVE_SUMMARY, 0.0);
      if (r.rate > 0.0) {
        int64_t rows = 0;
        if ((rc = dsm_flow_stress(e, r.ens, 0, 0, nullptr, &rows))) break;
        flow_rows.resize((size_t)rows * 9);
        if ((rc = dsm_flow_stress(e, r.ens, 0, rows, flow_rows.data(), &rows))) break;
        if (sm) {
          sm[0] = 2.0;
          sm[1] = (double)rows;
          if (rows > 0) std::copy(flow_rows.end() - 9, flow_rows.end(), sm + 2);
        }
      } else {
        int32_t np = 0;
        if ((rc = dsm_correlator(e, r.ens, 0, nullptr, nullptr, &np))) break;
        corr_lag.resize(np);
        corr_g.resize(np);
        if ((rc = dsm_correlator(e, r.ens, np, corr_lag.data(), corr_g.data(), &np))) break;
        if (sm) {
          sm[0] = 1.0;
          sm[1] = (double)np;
          for (int j = 0; j < 16 && j < np; ++j) sm[2 + j] = corr_g[j];
        }
      }
      if ((rc = dsm_remove_ensemble(e, r.ens))) break;
      used -= r.bytes;
      r.done = now();
      ++completed;
    }
    report->fetch_s += now() - tf;
    if (rc) break;
    active.swap(still);
  }
  const double t_end = now();
  dsm_destroy(e);
  if (rc) return rc;
  std::vector<double> lat;
  for (size_t i = 0; i < rq.size(); ++i) {
    const double l = rq[i].done - rq[i].arrival;
    lat.push_back(l);
    if (latency_s) latency_s[i] = l;
  }
  report->completed = completed;
  report->makespan_s = t_end;
  report->p50_s = quantile(lat, 0.50);
  report->p99_s = quantile(lat, 0.99);
  report->max_s = lat.empty() ? 0.0 : *std::max_element(lat.begin(), lat.end());
  double sum = 0.0;
  for (double l : lat) sum += l;
  report->mean_s = lat.empty() ? 0.0 : sum / (double)lat.size();
  report->jumps_per_s = t_end > 0.0 ? (double)report->jumps / t_end : 0.0;
  return DSM_OK;
}
