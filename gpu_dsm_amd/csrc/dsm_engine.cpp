// dsm_engine.cpp -- host side of the C ABI (include/dsm.h): ensembles, HBM arenas,
// staging windows, Z-binned batched launches and the correlator / flow-stress
// bookkeeping.  Compiled by hipcc for gfx950 together with dsm_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <map>
#include <string>
#include <vector>

#include "dsm.h"
#include "dsm_launch.h"
#include "dsm_types.h"

using namespace dsm;

namespace {

thread_local std::string g_err;

int fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  g_err = buf;
  return code;
}

#define HIPCHK(x)                                                                       \
  do {                                                                                  \
    hipError_t _e = (x);                                                                \
    if (_e != hipSuccess) return fail(DSM_ERR_DEVICE, "%s: %s", #x, hipGetErrorString(_e)); \
  } while (0)

constexpr int kNumBins = kNumGeoms;  // work bins = KMC chain geometries (capacity 64 << b)
// Work lists (cap entries each).  Sequential pass: list b feeds bin b, overflow to b+1.
// Concurrent pass: list 0 holds the pass's chains, lists 1..kNumBins the per-geometry
// split (the first launches), lists kNumBins+1.. the chains that outgrew their geometry
// there (the late launches: late list of bin b = kNumBins + b).
constexpr int kNumLists = 2 * kNumBins + 1;
constexpr int kSplitList = 1, kLateList = kNumBins;
constexpr int kMaxNk = 2047;         // Z <= Nk must fit the largest geometry after a creation
constexpr int kDefaultWindow = 256;
constexpr int kTotalsSlot = 4;       // d_counters: [0..3] advance_impl, [4..5] chain_totals
constexpr int kNumCounters = 6;
constexpr int kLdsTableFloats = 2816;  // per-wave LDS budget for rate tables (Nk <= ~200)

// ---------------------------------------------------------------- p_CD (FP32 host, as the reference)
// common/pcd_tau.h:32-63 and the device constants of common/gpu_random.cu:76-99.
struct Pcd {
  float g, alpha, tau_0, tau_max, tau_d, c1, cm1, rcm1c1, At, Adt, Bdt, normdt;
  float s_At, s_Ct, s_Dt, s_Adt, s_Bdt, s_Cdt, s_Ddt, s_taud_inv, s_one_minus_g;
};

void pcd_init(Pcd& p, float Be, float Nk) {
  const double z = (Nk + Be) / (Be + 1.0);
  p.g = 0.667f;
  if (Be != 1.0f) {
    p.alpha = (0.053f * logf(Be) + 0.31f) * powf((float)z, -0.012f * logf(Be) - 0.024f);
    p.tau_0 = 0.285f * powf(Be + 2.0f, 0.515f);
    p.tau_max = ((Nk < 2) ? p.tau_0 : 0.025f * powf(Be + 2.0f, 2.6f) * powf((float)z, 2.83f));
    p.tau_d = ((Nk < 2) ? p.tau_0 : 0.036f * powf(Be + 2.0f, 3.07f) * powf((float)(z - 1.0f), 3.02f));
  } else {
    p.alpha = 0.267096f - 0.375571f * expf(-0.0838237f * Nk);
    p.tau_0 = 0.460277f + 0.298913f * expf(-0.0705314f * Nk);
    p.tau_max = ((Nk < 4) ? p.tau_0 : 0.0156137f * powf(Nk, 3.18849f));
    p.tau_d = ((Nk < 4) ? p.tau_0 : 0.0740131f * powf(Nk, 3.18363f));
  }
  p.c1 = powf(p.tau_max, p.alpha) - powf(p.tau_0, p.alpha);
  p.cm1 = powf(p.tau_max, p.alpha - 1.0f) - powf(p.tau_0, p.alpha - 1.0f);
  p.rcm1c1 = ((p.c1 == 0.0f) ? (p.alpha - 1.0f) / p.alpha / p.tau_0 : (p.cm1 / p.c1));
  p.At = (1.0f - p.g);
  p.Adt = p.At * p.alpha / (p.alpha - 1.0f);
  p.Bdt = p.Adt * p.rcm1c1;
  p.normdt = p.Bdt + p.g / p.tau_d;
  p.s_taud_inv = 1.0f / p.tau_d;
  p.s_At = 1.0f * p.c1 / p.At;
  p.s_Dt = powf(p.tau_0, p.alpha);
  p.s_Ct = -1.0f / p.alpha;
  p.s_Adt = p.normdt * p.c1 / p.Adt;
  p.s_Bdt = p.Bdt / p.normdt;
  p.s_Cdt = -1.0f / (p.alpha - 1.0f);
  p.s_Ddt = powf(p.tau_0, p.alpha - 1.0f);
  p.s_one_minus_g = 1.0f - p.g;
}

double pow075(double x) { return std::sqrt(x) * std::sqrt(std::sqrt(x)); }

// ------------------------------------------------------------------ CD partner lengths
// Polydisperse CD background: the chain on the other side of an entanglement has a
// length drawn from the weight distribution of the melt (common/gpu_random.cu:62-76,
// 212-228, common/chain.cu:141-149).  The distribution is held as kPartnerK equiprobable
// quantiles (the Nk at probability (k+1/2)/K), each with its p_CD sampler constants;
// a draw uses the top kPartnerBits bits of a Philox word as the quantile index.

// Regularised lower incomplete gamma P(s, x): the power series of gamma(s, x) for
// x < s + 1, else 1 - Q from Legendre's continued fraction (modified Lentz).
double gamma_p(double s, double x) {
  if (x <= 0.0) return 0.0;
  const double lead = -x + s * std::log(x) - std::lgamma(s);
  if (x < s + 1.0) {
    double term = 1.0 / s, sum = term;
    for (int n = 1; n < 100000; ++n) {
      term *= x / (s + n);
      sum += term;
      if (std::fabs(term) < std::fabs(sum) * 1e-17) break;
    }
    return sum * std::exp(lead);
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - s, c = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 100000; ++i) {
    const double an = -i * (i - s);
    b += 2.0;
    d = an * d + b;
    if (std::fabs(d) < tiny) d = tiny;
    c = b + an / c;
    if (std::fabs(c) < tiny) c = tiny;
    d = 1.0 / d;
    const double del = d * c;
    h *= del;
    if (std::fabs(del - 1.0) < 1e-16) break;
  }
  return 1.0 - std::exp(lead) * h;
}

// Nk of the K quantiles of the partner weight distribution.  mode 1: GEX, weight
// fraction W(m) = P((a+1)/b, m^b) of the reduced mass m, Nk = m*mp/Mk
// (common/gamma.cpp:174-263, polydisp.dat); mode 2: the ensemble's own lengths, each
// distinct Nk weighted by Nk times its chain count.
bool partner_quantiles(int mode, const float gex[4], const std::vector<int>& nk_chain, std::vector<float>& out) {
  out.assign(kPartnerK, 0.0f);
  if (mode == 1) {
    const double a = gex[0], b = gex[1], mp = gex[2], Mk = gex[3];
    if (!(a > -1.0) || !(b > 0.0) || !(mp > 0.0) || !(Mk > 0.0)) return false;
    const double s = (a + 1.0) / b;
    for (int k = 0; k < kPartnerK; ++k) {
      const double y = (k + 0.5) / kPartnerK;
      double lo = 0.0, hi = 1.0;
      while (gamma_p(s, hi) < y) hi *= 2.0;
      for (int it = 0; it < 64; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (gamma_p(s, mid) < y) lo = mid; else hi = mid;
      }
      const double m = std::pow(0.5 * (lo + hi), 1.0 / b);
      out[k] = (float)(m * mp / Mk);
    }
    return true;
  }
  if (mode == 2) {
    std::map<int, double> w;
    for (int n : nk_chain) w[n] += (double)n;
    double tot = 0.0;
    for (auto& kv : w) tot += kv.second;
    auto it = w.begin();
    double cum = it->second;
    for (int k = 0; k < kPartnerK; ++k) {
      const double y = (k + 0.5) / kPartnerK * tot;
      while (cum < y && std::next(it) != w.end()) { ++it; cum += it->second; }
      out[k] = (float)it->first;
    }
    return true;
  }
  return false;
}

// Partner tables of one ensemble: for SD-created and initial links (f_t) the K
// quantiles above; for CD-created links (f_d_t) the same entries re-weighted by their
// mean destruction rate normdt (common/gamma.cpp:236-262).  Layout: [F consts 5K]
// [D consts 5K][F Nk K][D Nk K]; F consts s_At s_Ct s_Dt s_taud_inv s_one_minus_g,
// D consts s_Adt s_Bdt s_Cdt s_Ddt s_taud_inv.  Returns the CD creation prefactor
// <normdt>_F / beta (common/ensemble.cu:137-146).
float partner_tables(const std::vector<float>& nkF, float beta, std::vector<float>& t) {
  const int K = kPartnerK;
  t.assign((size_t)12 * K, 0.0f);
  std::vector<double> cum(K);
  std::vector<Pcd> pc(K);
  double tot = 0.0;
  for (int k = 0; k < K; ++k) {
    pcd_init(pc[k], beta, nkF[k]);
    tot += (double)pc[k].normdt;
    cum[k] = tot;
    float* f = &t[(size_t)5 * k];
    f[0] = pc[k].s_At; f[1] = pc[k].s_Ct; f[2] = pc[k].s_Dt; f[3] = pc[k].s_taud_inv; f[4] = pc[k].s_one_minus_g;
    t[(size_t)10 * K + k] = nkF[k];
  }
  int i = 0;
  for (int j = 0; j < K; ++j) {
    const double y = (j + 0.5) / K * tot;
    while (cum[i] < y && i + 1 < K) ++i;
    float* d = &t[(size_t)5 * K + (size_t)5 * j];
    d[0] = pc[i].s_Adt; d[1] = pc[i].s_Bdt; d[2] = pc[i].s_Cdt; d[3] = pc[i].s_Ddt; d[4] = pc[i].s_taud_inv;
    t[(size_t)11 * K + j] = nkF[i];
  }
  return (float)(tot / K) / beta;
}

// FP32 rate tables, indexed by the integer strand length N (see dsm_kernels.hip pair_rate).
void build_tables(std::vector<float>& t, int nmax, float beta, int cd, float c_create) {
  const int n = nmax + 2;
  const int fr = 4 * nmax + 4;
  t.assign((size_t)kNumTables * n + fr, 0.0f);
  float* SP = &t[0]; float* SM = SP + n; float* PP = SM + n; float* PM = PP + n; float* EC = PM + n;
  float* ED = EC + n; float* ED2 = ED + n; float* T3 = ED2 + n; float* CC = T3 + n; float* FR = CC + n;
  for (int N = 1; N < n; ++N) {
    const float fN = (float)N;
    SP[N] = 0.75f / (fN * (float)(N + 1));
    SM[N] = N >= 2 ? 0.75f / (fN * (float)(N - 1)) : 0.0f;
    PP[N] = (float)pow075((double)N / (double)(N + 1));
    PM[N] = N >= 2 ? (float)pow075((double)N / (double)(N - 1)) : 0.0f;
    EC[N] = 2.0f / (beta * (fN + 0.5f));
    ED[N] = 1.0f / (0.75f + 0.5f * fN);
    ED2[N] = 1.0f / (0.75f + (fN + 0.25f));
    T3[N] = 3.0f / fN;
    CC[N] = cd ? c_create * (float)(N - 1) : 0.0f;
  }
  for (int x = 1; x < fr; ++x) FR[x] = 4.0f / (float)x;
}

int bin_for(int nmax) {  // smallest geometry with capacity > nmax (room for one insertion)
  for (int b = 0; b < kNumBins; ++b)
    if (kmc_geom_capacity(b) > nmax) return b;
  return -1;
}

// strand slots per chain: Nk rounded up to whole 16-byte groups (aligned vector access)
inline int cap_of(int nk) { return (nk + 3) & ~3; }

// First-fit allocator over an index space [0, end) (chains, strand slots, staging
// floats): the ranges of removed ensembles are handed out again, so a long-running
// engine (serving) does not grow without bound.
struct RangeAlloc {
  std::map<int64_t, int64_t> free_;  // start -> length
  int64_t end = 0;
  int64_t take(int64_t n) {
    if (n <= 0) return end;
    for (auto it = free_.begin(); it != free_.end(); ++it) {
      if (it->second >= n) {
        const int64_t s0 = it->first, len = it->second;
        free_.erase(it);
        if (len > n) free_[s0 + n] = len - n;
        return s0;
      }
    }
    if (!free_.empty()) {  // a free range at the end is extended
      auto last = std::prev(free_.end());
      if (last->first + last->second == end) {
        const int64_t s0 = last->first;
        free_.erase(last);
        end = s0 + n;
        return s0;
      }
    }
    const int64_t s0 = end;
    end += n;
    return s0;
  }
  void give(int64_t s0, int64_t n) {
    if (n <= 0) return;
    auto it = free_.emplace(s0, n).first;
    if (it != free_.begin()) {
      auto prev = std::prev(it);
      if (prev->first + prev->second == it->first) {
        prev->second += it->second;
        free_.erase(it);
        it = prev;
      }
    }
    auto next = std::next(it);
    if (next != free_.end() && it->first + it->second == next->first) {
      it->second += next->second;
      free_.erase(next);
    }
  }
};

template <typename T>
int dev_alloc(T** p, size_t n) {
  *p = nullptr;
  if (n == 0) n = 1;
  hipError_t e = hipMalloc((void**)p, n * sizeof(T));
  if (e != hipSuccess) return fail(DSM_ERR_MEMORY, "hipMalloc(%zu bytes): %s", n * sizeof(T), hipGetErrorString(e));
  return DSM_OK;
}

// G(t) and MSD(t) both sample the equilibrium state every delta into a multi-tau correlator
inline bool eq_sampling(int mode) { return mode == kSampleEq || mode == kSampleMsd; }

struct Ens {
  bool live = false;
  dsm_ensemble_desc d{};
  std::vector<int> nk;
  int chain0 = 0, n = 0, nmax = 0;
  int64_t strand0 = 0, nstrand = 0;
  EnsParams P{};
  int bin_max = 0;      // largest geometry bin a chain of this ensemble can need
  double z_sum_est = 0.0;  // sum over its chains of the equilibrium mean Z, (Nk + beta) / (1 + beta)
  // sampling
  int ncomp = 0;        // floats per sample per chain (3 G(t), 8 flow)
  int window = 0;
  int64_t k_cur = 0;    // next sample index every healthy chain will record
  int64_t stage_floats = 0;
  int64_t last_k0 = 0, last_rows = 0;
  // correlator (G(t))
  int p = 0, m = 0, s = 0;
  double* dD = nullptr;
  double* dA = nullptr;
  double* dC = nullptr;
  std::vector<int> ins, macc;
  std::vector<int64_t> nins;
  // flow output rows (9 doubles per row)
  std::vector<double> flow_rows;
  int tab_slot = -1, pd_slot = -1;  // table arena slots (rate tables, partner tables)
};

}  // namespace

// Scheduling constants of the engine, set by dsm_set_option (and by the autotuner);
// none of them changes a result, only how the passes are laid out on the device.
struct EngineOptions {
  int plain_jumps = 1;                   // DSM_OPT_PLAIN_JUMPS
  int small_per_cu = 12;                 // DSM_OPT_SMALL_PER_CU
  int split_margin_div = 16;             // DSM_OPT_SPLIT_MARGIN
  int lds_table_floats = kLdsTableFloats;  // DSM_OPT_LDS_TABLE_FLOATS
  int concurrent = -1;                   // DSM_OPT_CONCURRENT
  int blocks_per_cu = 0;                 // DSM_OPT_BLOCKS_PER_CU
  int window = kDefaultWindow;           // DSM_OPT_WINDOW
  int graph = 0;                         // DSM_OPT_GRAPH
  int wide_pair = 1;                     // DSM_OPT_WIDE_PAIR
};

struct dsm_engine {
  int device = 0;
  EngineOptions opt;
  hipStream_t st = nullptr;
  hipStream_t xs[4] = {};  // xs[1..3]: the other streams of a concurrent pass (xs[0] stays null; st is stream 0)
  hipEvent_t ev_split = nullptr, ev_bin[4] = {};
  hipEvent_t cev[4] = {nullptr, nullptr, nullptr, nullptr};  // concurrent pass: start, row start/end, end
  hipEvent_t lev[2 * kNumBins] = {};                        // sequential pass: each launch's start/end
  int n_cu = 0;
  // chains
  int nch = 0;
  int64_t nstr = 0;
  ChainArrays ch{};
  // ensembles
  std::vector<Ens> ens;
  EnsParams* d_ens = nullptr;
  int d_ens_cap = 0;
  // Rate tables (kind 0) and polydisperse partner tables (kind 1) in one float arena.  A
  // slot is shared by every ensemble whose parameters give the same tables and counts its
  // users; a slot whose last user is removed keeps its contents (a later ensemble with
  // the same key takes it back as is) until a table of the same kind and size needs room.
  std::vector<float> tabs;
  float* d_tabs = nullptr;
  size_t d_tabs_cap = 0;
  struct TabSlot {
    int kind = 0;
    int nmax = 0, cd = 0, mode = 0;
    float beta = 0.0f, c_create = 0.0f;
    float gex[4] = {0, 0, 0, 0};
    std::vector<float> nkF;  // melt partners: the quantile lengths
    int64_t off = 0, size = 0;
    int refs = 0;
  };
  std::vector<TabSlot> tab_slots;
  std::vector<int64_t> stage_off;  // per ensemble
  int64_t stage_used = 0;
  RangeAlloc chain_space, strand_space, stage_space;
  float* d_stage = nullptr;
  int64_t d_stage_cap = 0;
  // work lists: [bin][nch] + counts
  int* d_work = nullptr;
  int d_work_cap = 0;
  int* d_cnt = nullptr;     // [kNumLists] list counts, then [kNumLists] queues
  unsigned long long* d_counters = nullptr;
  int* d_corr_cnt = nullptr;
  int d_corr_cnt_cap = 0;
  double* d_red = nullptr;
  int d_red_cap = 0;
  double* d_rows = nullptr;  // flow rows of one window, all ensembles
  int64_t d_rows_cap = 0;
  int* d_healthy = nullptr;
  double* d_part = nullptr;  // per-block partial sums of dsm_stress
  hipEvent_t evs[4] = {nullptr, nullptr, nullptr, nullptr};
  std::vector<int> host_work;
  // Captured window passes (DSM_OPT_GRAPH).  A graph holds device pointers and the
  // engine's options; `gen` counts their changes, and a graph is replayed only while
  // the generation it was captured in is current.  Its window parameters come from the
  // pinned block h_win, rewritten before every replay.
  uint64_t gen = 1;
  bool capturing = false;
  struct Captured {
    std::vector<int64_t> key;  // gen, then (ensemble, rows) per ensemble of the window
    hipGraphExec_t exec = nullptr;
    uint64_t used = 0;
  };
  std::vector<Captured> graphs;
  uint64_t graph_clock = 0;
  void* h_win = nullptr;
  size_t h_win_cap = 0;
};

namespace {

int set_device(dsm_engine* e) {
  HIPCHK(hipSetDevice(e->device));
  return DSM_OK;
}

// grow a device array of n_old elements to n_new, preserving contents
template <typename T>
int grow(T** p, int64_t n_old, int64_t n_new, hipStream_t st) {
  T* q = nullptr;
  int rc = dev_alloc(&q, (size_t)n_new);
  if (rc) return rc;
  if (*p && n_old > 0) HIPCHK(hipMemcpyAsync(q, *p, (size_t)n_old * sizeof(T), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  if (*p) HIPCHK(hipFree(*p));
  *p = q;
  return DSM_OK;
}

// capacity after growth to `need`: half again the current capacity, so that a stream of
// additions copies each element a bounded number of times, unless the extra would take
// more than a quarter of the free device memory (then exactly `need`)
int64_t grown_capacity(int64_t cur, int64_t need, size_t bytes_per) {
  if (need <= cur) return cur;
  int64_t cap = std::max<int64_t>(need, cur + cur / 2);
  size_t free_b = 0, total_b = 0;
  if (hipMemGetInfo(&free_b, &total_b) != hipSuccess || (size_t)(cap - need) * bytes_per > free_b / 4) cap = need;
  return cap;
}

constexpr size_t kChainBytes = 7 * 4 + 8 + 8 + 8 + 4 + 8 + 8 + 8 + 12;  // per-chain arrays below
constexpr size_t kStrandBytes = 5 * 4;                             // Qx Qy Qz N w

int grow_chains(dsm_engine* e, int n_need, int64_t s_need) {
  ChainArrays& c = e->ch;
  int rc = 0;
  e->gen++;  // chain arrays or work lists may move
  if (n_need > e->nch) {
    const int n0 = e->nch;
    const int n_new = (int)std::min<int64_t>(grown_capacity(n0, n_need, kChainBytes), INT32_MAX / 2);
    if ((rc = grow(&c.Z, n0, n_new, e->st)) || (rc = grow(&c.nk, n0, n_new, e->st)) ||
        (rc = grow(&c.ens, n0, n_new, e->st)) || (rc = grow(&c.cap, n0, n_new, e->st)) ||
        (rc = grow(&c.stall, n0, n_new, e->st)) || (rc = grow(&c.cid, n0, n_new, e->st)) ||
        (rc = grow(&c.local, n0, n_new, e->st)) || (rc = grow(&c.base, n0, n_new, e->st)) ||
        (rc = grow(&c.t, n0, n_new, e->st)) || (rc = grow(&c.ev, n0, n_new, e->st)) ||
        (rc = grow(&c.dt_prev, n0, n_new, e->st)) || (rc = grow(&c.next_sample, n0, n_new, e->st)) ||
        (rc = grow(&c.ev_stop, n0, n_new, e->st)) || (rc = grow(&c.zsum, n0, n_new, e->st)) ||
        (rc = grow(&c.R1, 3 * (int64_t)n0, 3 * (int64_t)n_new, e->st)))
      return rc;
    e->nch = n_new;
  }
  if (s_need > e->nstr) {
    const int64_t s0 = e->nstr;
    const int64_t s_new = grown_capacity(s0, s_need, kStrandBytes);
    if ((rc = grow(&c.Qx, s0, s_new, e->st)) || (rc = grow(&c.Qy, s0, s_new, e->st)) ||
        (rc = grow(&c.Qz, s0, s_new, e->st)) || (rc = grow(&c.N, s0, s_new, e->st)) ||
        (rc = grow(&c.w, s0, s_new, e->st)))
      return rc;
    e->nstr = s_new;
  }
  // bin 0 pads every live ensemble to whole quads; one spare list past the last bin
  int live = 1;
  for (const Ens& E : e->ens) live += E.live ? 1 : 0;
  const int need = e->nch + 4 * (live + 2);
  if (need > e->d_work_cap) {
    if (e->d_work) HIPCHK(hipFree(e->d_work));
    e->d_work = nullptr;
    e->d_work_cap = 0;
    const int cap = need + need / 8 + 1024;
    if ((rc = dev_alloc(&e->d_work, (size_t)kNumLists * cap))) return rc;
    e->d_work_cap = cap;
  }
  return DSM_OK;
}

int upload_ens(dsm_engine* e) {
  const int n = (int)e->ens.size();
  if (n > e->d_ens_cap) {
    e->gen++;
    if (e->d_ens) HIPCHK(hipFree(e->d_ens));
    int rc = dev_alloc(&e->d_ens, (size_t)std::max(n, 16));
    if (rc) return rc;
    e->d_ens_cap = std::max(n, 16);
  }
  std::vector<EnsParams> h(n);
  for (int i = 0; i < n; ++i) h[i] = e->ens[i].P;
  HIPCHK(hipMemcpyAsync(e->d_ens, h.data(), sizeof(EnsParams) * n, hipMemcpyHostToDevice, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  return DSM_OK;
}

// floats [off, off+n) of the table arena to the device (all of it when the device copy grows)
int upload_tabs(dsm_engine* e, int64_t off, int64_t n) {
  if (e->tabs.size() > e->d_tabs_cap) {
    e->gen++;
    if (e->d_tabs) HIPCHK(hipFree(e->d_tabs));
    e->d_tabs = nullptr;
    e->d_tabs_cap = 0;
    const size_t cap = std::max(e->tabs.size() * 2, (size_t)4096);
    int rc = dev_alloc(&e->d_tabs, cap);
    if (rc) return rc;
    e->d_tabs_cap = cap;
    off = 0;
    n = (int64_t)e->tabs.size();
  }
  HIPCHK(hipMemcpy(e->d_tabs + off, e->tabs.data() + off, sizeof(float) * n, hipMemcpyHostToDevice));
  return DSM_OK;
}

bool same_tables(const dsm_engine::TabSlot& a, const dsm_engine::TabSlot& b) {
  if (a.kind != b.kind || a.beta != b.beta) return false;
  if (a.kind == 0) return a.nmax == b.nmax && a.cd == b.cd && a.c_create == b.c_create;
  if (a.mode != b.mode) return false;
  if (a.mode == DSM_PARTNER_GEX) return std::equal(a.gex, a.gex + 4, b.gex);
  return a.nkF == b.nkF;
}

// slot holding the tables of `key` (one more user); `make` fills them when no slot has
// them: a free slot of the same kind and size is reused, else the arena grows
template <typename Make>
int take_tables(dsm_engine* e, const dsm_engine::TabSlot& key, Make make, int* slot) {
  for (size_t i = 0; i < e->tab_slots.size(); ++i)
    if (same_tables(e->tab_slots[i], key)) {
      e->tab_slots[i].refs++;
      *slot = (int)i;
      return DSM_OK;
    }
  std::vector<float> data;
  dsm_engine::TabSlot t = key;
  make(data, t);
  int idx = -1;
  for (size_t i = 0; i < e->tab_slots.size(); ++i)
    if (e->tab_slots[i].refs == 0 && e->tab_slots[i].kind == key.kind && e->tab_slots[i].size == (int64_t)data.size()) {
      idx = (int)i;
      break;
    }
  if (idx < 0) {
    if (e->tabs.size() + data.size() > (size_t)INT32_MAX) return fail(DSM_ERR_MEMORY, "table arena full");
    t.off = (int64_t)e->tabs.size();
    e->tabs.resize(e->tabs.size() + data.size());
    e->tab_slots.push_back(t);
    idx = (int)e->tab_slots.size() - 1;
  } else {
    t.off = e->tab_slots[idx].off;
  }
  t.size = (int64_t)data.size();
  t.refs = 1;
  std::copy(data.begin(), data.end(), e->tabs.begin() + t.off);
  e->tab_slots[idx] = t;
  *slot = idx;
  return upload_tabs(e, t.off, t.size);
}

int ensure_stage(dsm_engine* e) {
  if (e->stage_used > e->d_stage_cap) {  // keeps the rows of live ensembles (last_samples)
    e->gen++;
    const int64_t cap = grown_capacity(e->d_stage_cap, e->stage_used, sizeof(float));
    int rc = grow(&e->d_stage, e->d_stage_cap, cap, e->st);
    if (rc) return rc;
    e->d_stage_cap = cap;
  }
  return DSM_OK;
}

Ens* get_ens(dsm_engine* e, int id) {
  if (!e || id < 0 || id >= (int)e->ens.size() || !e->ens[id].live) return nullptr;
  return &e->ens[id];
}

// correlator memory, sized when the level count is known
int corr_alloc(dsm_engine* e, Ens& E, int s) {
  if (E.s == s && E.dD) return DSM_OK;
  e->gen++;
  if (E.dD) { (void)hipFree(E.dD); (void)hipFree(E.dA); (void)hipFree(E.dC); E.dD = E.dA = E.dC = nullptr; }
  E.s = s;
  const size_t n = (size_t)E.n;
  int rc;
  if ((rc = dev_alloc(&E.dD, (size_t)s * E.p * 3 * n)) || (rc = dev_alloc(&E.dA, (size_t)s * 3 * n)) ||
      (rc = dev_alloc(&E.dC, (size_t)s * E.p * n)))
    return rc;
  HIPCHK(hipMemsetAsync(E.dD, 0, sizeof(double) * s * E.p * 3 * n, e->st));
  HIPCHK(hipMemsetAsync(E.dA, 0, sizeof(double) * s * 3 * n, e->st));
  HIPCHK(hipMemsetAsync(E.dC, 0, sizeof(double) * s * E.p * n, e->st));
  E.ins.assign(s, 0);
  E.macc.assign(s, 0);
  E.nins.assign(s, 0);
  if (3 * s > e->d_corr_cnt_cap) {
    if (e->d_corr_cnt) HIPCHK(hipFree(e->d_corr_cnt));
    if ((rc = dev_alloc(&e->d_corr_cnt, (size_t)3 * s))) return rc;
    e->d_corr_cnt_cap = 3 * s;
  }
  if (s * E.p > e->d_red_cap) {
    if (e->d_red) HIPCHK(hipFree(e->d_red));
    if ((rc = dev_alloc(&e->d_red, (size_t)s * E.p))) return rc;
    e->d_red_cap = s * E.p;
  }
  HIPCHK(hipStreamSynchronize(e->st));
  return DSM_OK;
}

// host copy of the shared correlator counters after `rows` level-0 inserts
void corr_advance_counters(Ens& E, int64_t rows) {
  for (int64_t r = 0; r < rows; ++r) {
    int lvl = 0;
    for (;;) {
      if (lvl == E.s) break;
      E.macc[lvl]++;
      const bool push = E.macc[lvl] == E.m;
      if (push) E.macc[lvl] = 0;
      E.ins[lvl] = (E.ins[lvl] + 1 == E.p) ? 0 : E.ins[lvl] + 1;
      E.nins[lvl]++;
      if (!push) break;
      ++lvl;
    }
  }
}

struct Timer {
  std::chrono::steady_clock::time_point t0 = std::chrono::steady_clock::now();
  double ms() const {
    return std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
};

int device_blocks(dsm_engine* e, int geom, bool plain, int lds_floats, bool ext) {
  int nb = kmc_max_blocks_per_cu(geom, plain, lds_floats, ext);
  if (e->opt.blocks_per_cu > 0) nb = std::min(nb, e->opt.blocks_per_cu);
  if (nb <= 0) nb = 1;
  return nb * e->n_cu;
}

KmcArgs base_args(dsm_engine* e) {
  KmcArgs a{};
  a.ch = e->ch;
  a.ens = e->d_ens;
  a.tabs = e->d_tabs;
  a.stage = e->d_stage;
  a.use_ev_stop = 0;
  a.lds_floats = 0;
  return a;
}

// sums of the per-chain jump counters and Z sums over the given ensembles, in counter
// slots 4..5 (slots 0..3 hold advance_impl's totals at the start and end of a call)
int chain_totals(dsm_engine* e, const std::vector<int>& ids, unsigned long long out[2]) {
  unsigned long long* slot = e->d_counters + kTotalsSlot;
  HIPCHK(launch_fill_u32(slot, 0u, 4, e->st));
  for (int id : ids)
    HIPCHK(launch_sum_counters(e->ch.ev, e->ch.zsum, e->ens[id].chain0, e->ens[id].n, slot, e->st));
  HIPCHK(hipMemcpyAsync(out, slot, 2 * sizeof(unsigned long long), hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  return DSM_OK;
}

// Work of one KMC pass over the given ensembles: the bin-0 list holds every chain,
// each ensemble padded with -1 to whole quads (the row form runs four chains of one
// ensemble per wave); the wave-form bins receive the chains that outgrow a geometry.
struct Pass {
  int n0 = 0;  // entries of the bin-0 list
  int bin_max = 0;
  bool concurrent = false;  // batched passes only: split by Z, geometries side by side
  int n_ens = 0;
  int n_live = 0;           // chains of the pass
  int pending = 0;          // launch times recorded, not yet read: 1 sequential (ns launches), 2 concurrent
  int ns = 0;
  bool range = false;       // one ensemble, chains first .. first+n_live-1, no bin-0 list written
  int first = 0;
  KmcArgs a{};
};

// `batched`: a batched pass (a one-ensemble sequential one reads its chains by range and
// needs no bin-0 list)
int pass_setup(dsm_engine* e, const std::vector<int>& ids, bool use_ev_stop, Pass& ps, bool batched = false) {
  int n0 = 0, bin_max = 0, tab_max = 0, ext = 0;
  double z_est = 0.0, n_all = 0.0;
  for (int id : ids) {
    z_est += e->ens[id].z_sum_est;
    n_all += e->ens[id].n;
    ext |= (e->ens[id].P.sample_mode == kSampleMsd || e->ens[id].P.partner || e->ens[id].P.flow) ? 1 : 0;
    n0 += (e->ens[id].n + 3) & ~3;
    bin_max = std::max(bin_max, e->ens[id].bin_max);
    tab_max = std::max(tab_max, 9 * (e->ens[id].nmax + 2) + 4 * e->ens[id].nmax + 4);
  }
  ps.n0 = n0;
  ps.bin_max = bin_max;
  ps.n_ens = (int)ids.size();
  ps.n_live = (int)n_all;
  // geometries side by side when chains beyond the first wave-form geometry can occur
  // (their runs are long and few) or when many chains are expected above the row form's
  // 63 positions (the pair form then carries the bulk)
  ps.concurrent = e->opt.concurrent >= 0 ? (e->opt.concurrent > 0 && bin_max >= 1)
                                         : (bin_max >= 3 || (bin_max >= 1 && z_est >= 56.0 * n_all));
  if (n0 == 0) return DSM_OK;
  ps.range = batched && !ps.concurrent && ids.size() == 1;
  ps.first = e->ens[ids[0]].chain0;
  if (!ps.range) {
    HIPCHK(launch_fill_u32(e->d_work, 0xFFFFFFFFu, n0, e->st));
    int off = 0;
    for (int id : ids) {
      HIPCHK(launch_iota(e->d_work + off, e->ens[id].chain0, e->ens[id].n, e->st));
      off += (e->ens[id].n + 3) & ~3;
    }
  }
  ps.a = base_args(e);
  ps.a.use_ev_stop = use_ev_stop ? 1 : 0;
  // rate tables go to LDS (one copy per wave) when they fit the per-wave budget
  ps.a.lds_floats = tab_max <= e->opt.lds_table_floats ? tab_max : 0;
  ps.a.ext = ext;
  ps.a.plain_jumps = std::max(1, e->opt.plain_jumps);
  return DSM_OK;
}

// Chains that outgrow the row (or pair) form during a pass continue in the wave form
// (geometry 2, Z <= 256): these few chains form the pass's tail, and the wave form has
// the lowest per-jump latency.  Last geometry a pass runs them through:
constexpr int kLateFirst = 2;
int late_last(int bin_max) { return std::max(bin_max, kLateFirst); }

// Batched pass with the geometries side by side: the chains are split by their current
// Z into one list per geometry and every geometry runs at once, each on its own
// stream, the largest first (its few long chains take their slots before the bulk fills
// the device), so the latency-bound runs of long chains overlap the bulk instead of
// following it.  Chains that outgrow their geometry during this round go to late lists
// that a second round runs through the larger geometries (the first of them behind the
// merged row-form launch of bins 0 and 1, the others in order after the whole first
// round).  Every chain makes the same jumps to the same targets as in the sequential
// order.
int pass_bins_concurrent(dsm_engine* e, Pass& ps, dsm_run_stats* rs) {
  int* cnt = e->d_cnt;
  const size_t cap = (size_t)e->d_work_cap;
  auto list = [&](int l) { return e->d_work + (size_t)l * cap; };
  KmcArgs a = ps.a;
  auto bind = [&](int in, int out) {
    a.work = list(in);
    a.n_work = cnt + in;
    a.queue = cnt + kNumLists + in;
    a.overflow = list(out);
    a.n_overflow = cnt + out;
  };
  auto grid = [&](int b) { return device_blocks(e, b, false, a.lds_floats, a.ext != 0); };
  auto late_of = [&](int b) { return kLateList + (b < kLateFirst ? kLateFirst : b + 1); };  // bin b's overflow
  const bool timed = !e->capturing;  // a captured pass is timed as a whole, around its replay
  if (timed) HIPCHK(hipEventRecord(e->cev[0], e->st));
  HIPCHK(launch_fill_u32(cnt, 0u, 2 * kNumLists, e->st));
  HIPCHK(launch_split_geoms(list(0), ps.n0, e->ch.Z, e->ch.stall, list(kSplitList), (int)cap, cnt + kSplitList,
                            e->opt.split_margin_div, e->st));
  HIPCHK(hipEventRecord(e->ev_split, e->st));
  // first launches.  A process has four hardware queues on this device, so four streams.
  // The longest chains (G = 8: few chains, a latency-bound run that sets the window's
  // critical path) go first, on the engine stream right behind the split, so their
  // waves -- each leaves no room for another wave on its SIMD -- are dispatched before the other
  // queues, released by the split's event, fill the device (launched on a side stream
  // they waited for a free SIMD: 140-180 ms for a 68 ms run).  G = 4 (also latency-bound)
  // then on the second stream, the G = 1 bulk behind it, the merged bin-0/bin-1 row forms on
  // the third stream, the G = 2 bulk on the fourth.  The engine stream then runs second
  // launches of the G = 2 and G = 1 bins, which share their work queues, so whichever
  // stream is free first takes the remaining chains (a launch that finds its queue empty
  // exits at once).  The G = 2 runs, longer than the G = 1 ones, start with the pass.
  // Measured on the 262,144-chain melt, per 16 tau_K window: 111-113 ms; 118-123 ms with
  // the G = 1 bulk at the start and G = 2 behind G = 8; 128-136 ms with both bins in one
  // launch, G = 2 first.  The G = 4 waves (260 registers) find few SIMDs with room beside
  // the bulk's persistent waves and run their chains a few after one another (110 ms for
  // runs of 21 ms); holding the bulk streams back 100 us lets them all start at once and
  // end at 21 ms, but the bulk then runs at a lower occupancy meanwhile: 114 ms.
  auto stream = [&](int k) { return k == 0 ? e->st : e->xs[k]; };  // xs[0] is not a stream
  // bin 2 (128 <= Z < 240 at the split) runs two chains per wave in the two-group pair
  // form unless DSM_OPT_WIDE_PAIR is 0 (a third fewer VALU instructions per jump; on the
  // 262,144-chain melt 93.6-94.2 ms per window against 95.0-96.7 ms in the G = 1 wave
  // form, since the strand lengths left the registers and with them most of its spills);
  // the late round keeps one chain per wave.  Only one-ensemble passes: a wave of two
  // chains runs one ensemble at a time, and in a serving mix of many ensembles its
  // segments would wait for each other (the C5 serving stream at horizon 60: 6.2e8
  // jumps/s with the two-chain form, 8.4e8 with one chain per wave; at horizon 40 the
  // p50 latency 5.6 s against 0.38 s).  The same switch puts bin 1 (64 <= Z < 128) in the
  // row form with two register groups, four chains per wave, where the pair form runs two
  // (one-ensemble 262,144-chain melt: 80-82 ms per window against 89.7-90.3 ms; the C5
  // stream at horizon 60: 7.9e8 jumps/s against 8.4e8 in the pair form)
  const bool wide = e->opt.wide_pair && ps.n_ens == 1;
  auto first = [&](int b, int k) -> hipError_t {
    if (b > ps.bin_max) return hipSuccess;
    bind(kSplitList + b, late_of(b));
    ++rs->kmc_launches;
    const int geom = (b == 2 && wide) ? kGeomPair2 : b;
    return launch_kmc(a, geom, false, grid(geom), stream(k));
  };
  HIPCHK(first(5, 0));
  for (int k = 1; k < 4; ++k) HIPCHK(hipStreamWaitEvent(e->xs[k], e->ev_split, 0));
  HIPCHK(first(4, 1));
  HIPCHK(first(2, 1));
  HIPCHK(first(3, 3));
  bind(kSplitList + 1, late_of(1));
  const KmcArgs pair_args = a;
  bind(kSplitList, late_of(0));
  int nb = kmc_rowpair_occupancy(a.lds_floats, a.ext != 0, wide);
  if (e->opt.blocks_per_cu > 0) nb = std::min(nb, e->opt.blocks_per_cu);
  if (timed) HIPCHK(hipEventRecord(e->cev[1], e->xs[2]));
  HIPCHK(launch_kmc_rowpair(a, pair_args, wide, std::max(nb, 1) * e->n_cu, e->xs[2]));
  if (timed) HIPCHK(hipEventRecord(e->cev[2], e->xs[2]));
  // bins 0 and 1 are the only feeders of the first late list, complete when
  // their launch ends: its run follows on the same stream, beside the other geometries'
  // bulk, instead of after the whole first round (118-123 ms per window against 123-126
  // with the earlier stream layout)
  bind(kLateList + kLateFirst, late_of(kLateFirst));
  HIPCHK(launch_kmc(a, kLateFirst, false, grid(kLateFirst), e->xs[2]));
  rs->kmc_launches += 2;
  HIPCHK(first(3, 0));
  HIPCHK(first(2, 0));
  for (int k = 1; k < 4; ++k) HIPCHK(hipEventRecord(e->ev_bin[k], e->xs[k]));
  for (int k = 1; k < 4; ++k) HIPCHK(hipStreamWaitEvent(e->st, e->ev_bin[k], 0));
  // late launches: chains that outgrew their geometry, through the larger ones in order
  for (int b = kLateFirst + 1; b <= late_last(ps.bin_max); ++b) {
    bind(kLateList + b, late_of(b));
    HIPCHK(launch_kmc(a, b, false, grid(b), e->st));
  }
  if (timed) HIPCHK(hipEventRecord(e->cev[3], e->st));
  ps.pending = timed ? 2 : 0;  // read by pass_times after the caller's next synchronisation
  rs->kmc_launches += late_last(ps.bin_max) - kLateFirst;
  rs->first_launches++;
  return DSM_OK;
}

// One pass through the geometry bins in order: the row form, then (for chains that
// outgrow it) the wave form and the larger geometries; list b feeds bin b.  Batched:
// each chain runs to its targets; plain: each chain makes one jump.  Times every launch
// when `timed` (one host wait at the end).
int pass_bins(dsm_engine* e, Pass& ps, bool plain, bool timed, dsm_run_stats* rs) {
  if (!plain && ps.concurrent) return pass_bins_concurrent(e, ps, rs);
  int* cnt = e->d_cnt;
  HIPCHK(launch_init_counts(cnt, 2 * kNumLists + 17, ps.n0, e->st));  // lists, queues, order scratch
  KmcArgs a = ps.a;
  int seq[kNumBins], ns = 0;
  seq[ns++] = 0;
  if (ps.bin_max >= 1)
    for (int b = kLateFirst; b <= late_last(ps.bin_max); ++b) seq[ns++] = b;
  // batched, one ensemble: the row form takes its chains in decreasing Z, so the chains
  // started last, which set the end of the launch, are the shortest (several ensembles
  // keep their order: a wave runs one ensemble at a time)
  // A pass of few chains -- at most one wave-form wave per chain at three waves per SIMD
  // -- is latency-bound: it runs in the wave form, whose per-jump latency is the lowest
  // (1.79 us per jump against 2.84 us per row-form step of four chains).  Measured:
  // 1000 chains of Nk 25 +15%, of Nk 100 +92%; 4096 chains of Nk 25 -14%.
  const bool small = !plain && ps.n_live <= e->opt.small_per_cu * e->n_cu;
  const bool zorder = !plain && (ps.n_ens == 1 || small);
  constexpr int kZList = kNumLists - 1;
  if (zorder)
    HIPCHK(launch_z_order(ps.range ? nullptr : e->d_work, ps.first, ps.range ? ps.n_live : ps.n0, e->ch.Z,
                          e->ch.stall, cnt + 2 * kNumLists, e->d_work + (size_t)kZList * e->d_work_cap,
                          cnt + kZList, e->st));
  if (small) {  // the compact ordered list feeds the wave form directly
    ns = 0;
    for (int b = kLateFirst; b <= late_last(ps.bin_max); ++b) seq[ns++] = b;
  }
  for (int k = 0; k < ns; ++k) {
    const int b = seq[k], nb = k + 1 < ns ? seq[k + 1] : b + 1;
    const int in = (k == 0 && zorder) ? kZList : b;
    a.work = e->d_work + (size_t)in * e->d_work_cap;
    a.n_work = cnt + in;
    a.queue = cnt + kNumLists + in;
    a.overflow = e->d_work + (size_t)nb * e->d_work_cap;
    a.n_overflow = cnt + nb;
    const int grid = device_blocks(e, b, plain, a.lds_floats, a.ext != 0);
    if (timed) HIPCHK(hipEventRecord(e->lev[2 * k], e->st));
    HIPCHK(launch_kmc(a, b, plain, grid, e->st));
    if (timed) HIPCHK(hipEventRecord(e->lev[2 * k + 1], e->st));
    rs->kmc_launches++;
    if (k == 0) rs->first_launches++;
  }
  if (timed) {
    ps.pending = 1;  // read by pass_times after the caller's next synchronisation
    ps.ns = ns;
  }
  return DSM_OK;
}

// Launch times of a timed pass, read once its events have completed (waits for them
// if the caller has not synchronised yet).
int pass_times(dsm_engine* e, Pass& ps, dsm_run_stats* rs) {
  float ms = 0.0f;
  if (ps.pending == 1) {
    HIPCHK(hipEventSynchronize(e->lev[2 * ps.ns - 1]));
    for (int k = 0; k < ps.ns; ++k) {
      HIPCHK(hipEventElapsedTime(&ms, e->lev[2 * k], e->lev[2 * k + 1]));
      rs->kmc_ms += ms;
      if (k == 0) rs->first_kmc_ms += ms;
    }
  } else if (ps.pending == 2) {
    HIPCHK(hipEventSynchronize(e->cev[3]));
    HIPCHK(hipEventElapsedTime(&ms, e->cev[0], e->cev[3]));
    rs->kmc_ms += ms;
    HIPCHK(hipEventElapsedTime(&ms, e->cev[1], e->cev[2]));
    rs->first_kmc_ms += ms;
  }
  ps.pending = 0;
  return DSM_OK;
}

// One KMC pass over every chain of the given ensembles until each reaches its
// ensemble's P.t_end (batched: persistent geometry-binned launches; plain: one jump
// per chain per pass until no chain moves).  P must already be uploaded.
// With `later`, a batched pass's launch times are left in *later for pass_times after
// the caller's own synchronisation (no host wait between the pass and what follows it).
// The run flag is not polled here: a pass always completes its window, whose samples
// the caller then consumes, so an aborted call leaves a state that continues exactly.
int kmc_pass(dsm_engine* e, const std::vector<int>& ids, int mode, bool use_ev_stop, dsm_run_stats* rs,
             Pass* later = nullptr) {
  Pass ps;
  int rc = pass_setup(e, ids, use_ev_stop, ps, mode != DSM_MODE_PLAIN);
  if (rc || ps.n0 == 0) return rc;
  if (mode == DSM_MODE_PLAIN) {
    const int batch = 16;
    for (;;) {
      unsigned long long before[2], after[2];
      if ((rc = chain_totals(e, ids, before))) return rc;
      HIPCHK(hipEventRecord(e->evs[2], e->st));
      for (int k = 0; k < batch; ++k)
        if ((rc = pass_bins(e, ps, true, false, rs))) return rc;
      HIPCHK(hipEventRecord(e->evs[3], e->st));
      if ((rc = chain_totals(e, ids, after))) return rc;
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, e->evs[2], e->evs[3]));
      rs->kmc_ms += ms;
      rs->first_kmc_ms += ms;
      if (after[0] == before[0]) break;
    }
    return DSM_OK;
  }
  if ((rc = pass_bins(e, ps, false, !e->capturing, rs))) return rc;
  if (later) *later = ps;
  else if ((rc = pass_times(e, ps, rs))) return rc;
  return DSM_OK;
}

int finish_stats(dsm_engine* e, const std::vector<int>& ids, const unsigned long long before[2],
                 dsm_run_stats* rs, const Timer& T) {
  unsigned long long c[2];
  int rc = chain_totals(e, ids, c);
  if (rc) return rc;
  rs->events = c[0] - before[0];
  rs->z_sum = c[1] - before[1];
  rs->total_ms = T.ms();
  return DSM_OK;
}

// The device work of one window after its parameters are set: staging fills, the KMC
// pass and the sample consumers (correlator update, flow-row reduction).  Parameters come
// from the host (eager: uploaded) or, in a captured pass, from the pinned block h_win
// through stage_window_kernel at the head of the graph.
int window_launches(dsm_engine* e, const std::vector<int>& act, const std::vector<int64_t>& cnt_off,
                    const std::vector<int64_t>& row_off, const std::vector<int>& cnt, int mode, dsm_run_stats* rs,
                    Pass* timed) {
  int rc;
  if (e->capturing) {
    const WinParam* wp = (const WinParam*)e->h_win;
    HIPCHK(launch_stage_window(wp, (int)act.size(), e->d_ens, (const int*)(wp + act.size()), e->d_corr_cnt,
                               (int)cnt.size(), e->st));
  } else {
    if ((rc = upload_ens(e))) return rc;
    if (!cnt.empty())
      HIPCHK(hipMemcpyAsync(e->d_corr_cnt, cnt.data(), sizeof(int) * cnt.size(), hipMemcpyHostToDevice, e->st));
  }
  for (int id : act) {
    Ens& E = e->ens[id];
    if (E.P.k_end > E.P.k0)
      HIPCHK(launch_fill_u32(e->d_stage + E.P.stage_off, 0u, (E.P.k_end - E.P.k0) * E.n * E.ncomp, e->st));
  }
  if ((rc = kmc_pass(e, act, mode, false, rs, timed))) return rc;
  for (size_t q = 0; q < act.size(); ++q) {
    Ens& E = e->ens[act[q]];
    const int rows = (int)(E.P.k_end - E.P.k0);
    if (rows <= 0) continue;
    if (eq_sampling(E.P.sample_mode))
      HIPCHK(launch_corr_update(e->d_stage + E.P.stage_off, E.dD, E.dA, E.dC, e->ch.stall, E.chain0, E.n, E.p, E.m,
                                E.s, rows, e->d_corr_cnt + cnt_off[q], E.P.sample_mode == kSampleMsd ? 1 : 0, e->st));
    else
      HIPCHK(launch_flow_reduce(e->d_stage + E.P.stage_off, E.n, rows, e->d_rows + row_off[q], e->st));
  }
  return DSM_OK;
}

// The window's launches captured into a graph (stream capture of the engine stream and the
// streams a concurrent pass forks to), instantiated for replay.
int capture_window(dsm_engine* e, const std::vector<int>& act, const std::vector<int64_t>& cnt_off,
                   const std::vector<int64_t>& row_off, const std::vector<int>& cnt, dsm_run_stats* rs,
                   hipGraphExec_t* exec) {
  *exec = nullptr;
  (void)rs;  // nothing runs during the capture: the replay is counted by the caller
  dsm_run_stats none{};
  HIPCHK(hipStreamBeginCapture(e->st, hipStreamCaptureModeThreadLocal));
  e->capturing = true;
  const int rc = window_launches(e, act, cnt_off, row_off, cnt, DSM_MODE_BATCHED, &none, nullptr);
  e->capturing = false;
  hipGraph_t g = nullptr;
  const hipError_t ce = hipStreamEndCapture(e->st, &g);
  if (rc) {
    if (g) (void)hipGraphDestroy(g);
    return rc;
  }
  if (ce != hipSuccess || !g) return fail(DSM_ERR_DEVICE, "stream capture of a window failed: %s", hipGetErrorString(ce));
  const hipError_t ie = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  if (ie != hipSuccess) return fail(DSM_ERR_DEVICE, "graph instantiation failed: %s", hipGetErrorString(ie));
  return DSM_OK;
}

void drop_graphs(dsm_engine* e) {
  for (auto& c : e->graphs)
    if (c.exec) (void)hipGraphExecDestroy(c.exec);
  e->graphs.clear();
}

// A captured window of this shape, captured now if there is none (at most kMaxGraphs kept).
constexpr size_t kMaxGraphs = 16;
int window_graph(dsm_engine* e, const std::vector<int64_t>& key, const std::vector<int>& act,
                 const std::vector<int64_t>& cnt_off, const std::vector<int64_t>& row_off, const std::vector<int>& cnt,
                 dsm_run_stats* rs, hipGraphExec_t* exec) {
  for (auto& c : e->graphs)
    if (c.key == key) {
      c.used = ++e->graph_clock;
      *exec = c.exec;
      return DSM_OK;
    }
  // graphs of an older generation can never be replayed again
  for (size_t i = 0; i < e->graphs.size();) {
    if (e->graphs[i].key[0] != key[0]) {
      (void)hipGraphExecDestroy(e->graphs[i].exec);
      e->graphs.erase(e->graphs.begin() + (long)i);
    } else {
      ++i;
    }
  }
  if (e->graphs.size() >= kMaxGraphs) {
    auto lru = std::min_element(e->graphs.begin(), e->graphs.end(),
                                [](const dsm_engine::Captured& a, const dsm_engine::Captured& b) { return a.used < b.used; });
    (void)hipGraphExecDestroy(lru->exec);
    e->graphs.erase(lru);
  }
  dsm_engine::Captured c;
  c.key = key;
  c.used = ++e->graph_clock;
  int rc = capture_window(e, act, cnt_off, row_off, cnt, rs, &c.exec);
  if (rc) return rc;
  e->graphs.push_back(c);
  *exec = c.exec;
  return DSM_OK;
}

int advance_impl(dsm_engine* e, const std::vector<int>& ids, const std::vector<double>& t_end, int mode,
                 volatile const int32_t* run_flag, volatile int32_t* progress, dsm_run_stats* out) {
  Timer T;
  dsm_run_stats rs{};
  int rc = set_device(e);
  if (rc) return rc;
  // jump and Z totals before (d_counters[0..1]) and after (d_counters[2..3]) the call,
  // summed on the device: the "after" sums and their copy ride on the last window's
  // synchronisation, so the totals cost no host round trip of their own
  unsigned long long tot[4] = {0, 0, 0, 0};
  bool have_tot = false;
  HIPCHK(launch_fill_u32(e->d_counters, 0u, 8, e->st));
  for (int id : ids)
    HIPCHK(launch_sum_counters(e->ch.ev, e->ch.zsum, e->ens[id].chain0, e->ens[id].n, e->d_counters, e->st));
  auto aborted = [&]() {
    unsigned long long before[2] = {0, 0};
    if (hipMemcpy(before, e->d_counters, sizeof(before), hipMemcpyDeviceToHost) == hipSuccess)
      finish_stats(e, ids, before, &rs, T);
    if (out) *out = rs;
    return DSM_ABORTED;
  };
  // per-ensemble sample targets
  std::vector<int64_t> k_target(ids.size(), 0);
  for (size_t q = 0; q < ids.size(); ++q) {
    Ens& E = e->ens[ids[q]];
    const double te = t_end[q];
    if (eq_sampling(E.P.sample_mode)) {
      k_target[q] = (int64_t)std::ceil(te / E.P.delta - 1e-12);
      if (E.s == 0) {
        int s = E.d.corr_levels;
        if (s <= 0) s = (int)std::ceil(std::log(te / E.P.delta / E.p) / std::log((double)E.m)) + 1;
        if (s < 1) s = 1;
        if (s > 40) return fail(DSM_ERR_ARG, "correlator needs %d levels (max 40)", s);
        if ((rc = corr_alloc(e, E, s))) return rc;
      }
    } else if (E.P.sample_mode == kSampleFlow) {
      k_target[q] = (int64_t)std::ceil(te / E.P.delta - 1e-12) + 1;
    }
  }
  const double t0_prog = 0.0;
  (void)t0_prog;
  std::vector<char> reached(ids.size(), 0);
  for (;;) {
    // next window: each ensemble advances to its window end, the last one to t_end
    std::vector<int> act;
    for (size_t q = 0; q < ids.size(); ++q) {
      if (reached[q]) continue;
      Ens& E = e->ens[ids[q]];
      if (E.P.sample_mode == kSampleNone || E.k_cur >= k_target[q]) {
        E.P.t_end = t_end[q];
        E.P.k0 = E.k_cur;
        E.P.k_end = E.k_cur;
        act.push_back(ids[q]);
        reached[q] = 1;
        continue;
      }
      const int64_t k_end = std::min(E.k_cur + E.window, k_target[q]);
      E.P.k0 = E.k_cur;
      E.P.k_end = k_end;
      if (eq_sampling(E.P.sample_mode)) E.P.t_end = std::min((double)k_end * E.P.delta, t_end[q]);
      else E.P.t_end = (double)(k_end - 1) * E.P.delta;  // flow output k_end-1 is at this time
      if (k_end == k_target[q]) reached[q] = 1;
      act.push_back(ids[q]);
    }
    if (act.empty()) break;
    const bool last = std::all_of(reached.begin(), reached.end(), [](char r) { return r != 0; });
    // the window's sample consumers: level counters set before the pass
    std::vector<int> cnt;
    std::vector<int64_t> cnt_off(act.size(), -1), row_off(act.size(), -1);
    int64_t flow_doubles = 0;
    for (size_t q = 0; q < act.size(); ++q) {
      const Ens& E = e->ens[act[q]];
      if (E.P.k_end <= E.P.k0) continue;
      if (eq_sampling(E.P.sample_mode)) {
        cnt_off[q] = (int64_t)cnt.size();
        for (int k = 0; k < E.s; ++k) cnt.push_back(E.ins[k]);
        for (int k = 0; k < E.s; ++k) cnt.push_back(E.macc[k]);
        for (int k = 0; k < E.s; ++k) cnt.push_back((int)std::min<int64_t>(E.nins[k], E.p));
      } else {
        row_off[q] = flow_doubles;
        flow_doubles += (E.P.k_end - E.P.k0) * 9;
      }
    }
    if ((int)cnt.size() > e->d_corr_cnt_cap) {
      e->gen++;
      if (e->d_corr_cnt) HIPCHK(hipFree(e->d_corr_cnt));
      if ((rc = dev_alloc(&e->d_corr_cnt, cnt.size()))) return rc;
      e->d_corr_cnt_cap = (int)cnt.size();
    }
    if (flow_doubles > e->d_rows_cap) {
      e->gen++;
      if (e->d_rows) HIPCHK(hipFree(e->d_rows));
      if ((rc = dev_alloc(&e->d_rows, (size_t)flow_doubles))) return rc;
      e->d_rows_cap = flow_doubles;
    }
    Pass timed;
    bool graph_timed = false;
    if (e->opt.graph && mode == DSM_MODE_BATCHED) {
      // captured pass: window parameters through the pinned block, the launches replayed
      const size_t need = sizeof(WinParam) * act.size() + sizeof(int) * cnt.size();
      if (need > e->h_win_cap) {
        e->gen++;
        if (e->h_win) HIPCHK(hipHostFree(e->h_win));
        e->h_win = nullptr;
        e->h_win_cap = 0;
        const size_t cap = std::max<size_t>(4096, 2 * need);
        HIPCHK(hipHostMalloc(&e->h_win, cap, hipHostMallocCoherent));
        e->h_win_cap = cap;
      }
      WinParam* wp = (WinParam*)e->h_win;
      for (size_t q = 0; q < act.size(); ++q) {
        const Ens& E = e->ens[act[q]];
        wp[q] = WinParam{act[q], 0, E.P.t_end, E.P.k0, E.P.k_end};
      }
      if (!cnt.empty()) std::memcpy(wp + act.size(), cnt.data(), sizeof(int) * cnt.size());
      std::vector<int64_t> key{(int64_t)e->gen};
      for (int id : act) {
        key.push_back(id);
        key.push_back(e->ens[id].P.k_end - e->ens[id].P.k0);
      }
      hipGraphExec_t exec = nullptr;
      if ((rc = window_graph(e, key, act, cnt_off, row_off, cnt, &rs, &exec))) return rc;
      HIPCHK(hipEventRecord(e->evs[0], e->st));
      HIPCHK(hipGraphLaunch(exec, e->st));
      HIPCHK(hipEventRecord(e->evs[1], e->st));
      graph_timed = true;
      rs.kmc_launches++;
      rs.first_launches++;
    } else {
      if ((rc = window_launches(e, act, cnt_off, row_off, cnt, mode, &rs, &timed))) return rc;
    }
    // consume the window's samples: every reduction is queued, then one synchronisation
    std::vector<double> flow_host((size_t)flow_doubles);
    if (flow_doubles)
      HIPCHK(hipMemcpyAsync(flow_host.data(), e->d_rows, sizeof(double) * flow_doubles, hipMemcpyDeviceToHost,
                            e->st));
    if (last) {
      for (int id : ids)
        HIPCHK(launch_sum_counters(e->ch.ev, e->ch.zsum, e->ens[id].chain0, e->ens[id].n, e->d_counters + 2,
                                   e->st));
      HIPCHK(hipMemcpyAsync(tot, e->d_counters, sizeof(tot), hipMemcpyDeviceToHost, e->st));
      have_tot = true;
    }
    HIPCHK(hipStreamSynchronize(e->st));
    if ((rc = pass_times(e, timed, &rs))) return rc;
    if (graph_timed) {
      float ms = 0.0f;
      HIPCHK(hipEventElapsedTime(&ms, e->evs[0], e->evs[1]));
      rs.kmc_ms += ms;
      rs.first_kmc_ms += ms;
    }
    for (size_t q = 0; q < act.size(); ++q) {
      Ens& E = e->ens[act[q]];
      const int64_t rows = E.P.k_end - E.P.k0;
      if (rows > 0) {
        E.last_k0 = E.P.k0;
        E.last_rows = rows;
        if (eq_sampling(E.P.sample_mode)) {
          corr_advance_counters(E, rows);
        } else {
          const double* src = flow_host.data() + row_off[q];
          E.flow_rows.insert(E.flow_rows.end(), src, src + rows * 9);
        }
      }
      E.k_cur = E.P.k_end;
    }
    if (progress && !ids.empty()) {
      Ens& E = e->ens[ids[0]];
      const double frac = k_target[0] > 0 ? (double)E.k_cur / (double)k_target[0] : 1.0;
      *progress = (int32_t)(100.0 * frac);
    }
    if (run_flag && *run_flag == 0 && !last) return aborted();
  }
  if (have_tot) {
    rs.events = tot[2] - tot[0];
    rs.z_sum = tot[3] - tot[1];
    rs.total_ms = T.ms();
  } else {
    unsigned long long before[2] = {0, 0};
    HIPCHK(hipMemcpy(before, e->d_counters, sizeof(before), hipMemcpyDeviceToHost));
    if ((rc = finish_stats(e, ids, before, &rs, T))) return rc;
  }
  if (out) *out = rs;
  return DSM_OK;
}

}  // namespace

// ================================================================== C ABI
extern "C" {

const char* dsm_last_error(void) { return g_err.c_str(); }

int dsm_mem_info(int32_t device_id, size_t* free_bytes, size_t* total_bytes) {
  HIPCHK(hipSetDevice(device_id));
  size_t f = 0, t = 0;
  HIPCHK(hipMemGetInfo(&f, &t));
  if (free_bytes) *free_bytes = f;
  if (total_bytes) *total_bytes = t;
  return DSM_OK;
}

int dsm_create(int32_t device_id, dsm_engine** out) {
  if (!out) return fail(DSM_ERR_ARG, "null output pointer");
  *out = nullptr;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return fail(DSM_ERR_DEVICE, "no HIP device found");
  if (device_id < 0 || device_id >= ndev) return fail(DSM_ERR_DEVICE, "device %d out of range (%d devices)", device_id, ndev);
  HIPCHK(hipSetDevice(device_id));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device_id));
  dsm_engine* e = new dsm_engine();
  e->device = device_id;
  e->n_cu = prop.multiProcessorCount;
  bool streams_ok = hipStreamCreateWithFlags(&e->st, hipStreamNonBlocking) == hipSuccess;
  for (int k = 1; k < 4; ++k)
    streams_ok = streams_ok && hipStreamCreateWithFlags(&e->xs[k], hipStreamNonBlocking) == hipSuccess;
  if (!streams_ok) {
    dsm_destroy(e);
    return fail(DSM_ERR_DEVICE, "stream creation failed");
  }
  for (auto& ev : e->evs) HIPCHK(hipEventCreate(&ev));
  for (auto& ev : e->cev) HIPCHK(hipEventCreate(&ev));
  for (auto& ev : e->lev) HIPCHK(hipEventCreate(&ev));
  HIPCHK(hipEventCreateWithFlags(&e->ev_split, hipEventDisableTiming));
  for (auto& ev : e->ev_bin) HIPCHK(hipEventCreateWithFlags(&ev, hipEventDisableTiming));
  int rc;
  if ((rc = dev_alloc(&e->d_cnt, 2 * kNumLists + 32)) || (rc = dev_alloc(&e->d_counters, kNumCounters)) ||
      (rc = dev_alloc(&e->d_healthy, 1)) || (rc = dev_alloc(&e->d_part, (size_t)kStressBlocks * 9))) {
    dsm_destroy(e);
    return rc;
  }
  // verify the code object runs on this device
  hipError_t le = launch_fill_u32(e->d_counters, 0u, 8, e->st);
  if (le != hipSuccess || hipStreamSynchronize(e->st) != hipSuccess) {
    dsm_destroy(e);
    return fail(DSM_ERR_NOKERNEL, "gfx950 code object did not launch on device %d (%s)", device_id, prop.gcnArchName);
  }
  *out = e;
  return DSM_OK;
}

int dsm_destroy(dsm_engine* e) {
  if (!e) return DSM_OK;
  (void)hipSetDevice(e->device);
  if (e->st) (void)hipStreamSynchronize(e->st);
  for (auto& x : e->xs)
    if (x) (void)hipStreamSynchronize(x);
  drop_graphs(e);
  if (e->h_win) (void)hipHostFree(e->h_win);
  for (auto& E : e->ens) {
    if (E.dD) (void)hipFree(E.dD);
    if (E.dA) (void)hipFree(E.dA);
    if (E.dC) (void)hipFree(E.dC);
  }
  ChainArrays& c = e->ch;
  void* ptrs[] = {c.Z, c.nk, c.ens, c.cap, c.stall, c.cid, c.local, c.base, c.t, c.ev, c.dt_prev, c.next_sample, c.ev_stop, c.zsum, c.R1,
                  c.Qx, c.Qy, c.Qz, c.N, c.w, e->d_ens, e->d_tabs, e->d_stage, e->d_work, e->d_cnt, e->d_counters,
                  e->d_corr_cnt, e->d_red, e->d_rows, e->d_healthy, e->d_part};
  for (void* p : ptrs)
    if (p) (void)hipFree(p);
  for (auto& ev : e->evs)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& ev : e->cev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& ev : e->lev)
    if (ev) (void)hipEventDestroy(ev);
  for (auto& ev : e->ev_bin)
    if (ev) (void)hipEventDestroy(ev);
  if (e->ev_split) (void)hipEventDestroy(e->ev_split);
  for (auto& x : e->xs)
    if (x) (void)hipStreamDestroy(x);
  if (e->st) (void)hipStreamDestroy(e->st);
  delete e;
  return DSM_OK;
}

int dsm_add_ensemble(dsm_engine* e, const dsm_ensemble_desc* d, int32_t* ens_id) {
  if (!e || !d) return fail(DSM_ERR_ARG, "null argument");
  if (d->n_chains <= 0) return fail(DSM_ERR_ARG, "n_chains must be positive");
  if (!(d->beta > 0.0f)) return fail(DSM_ERR_ARG, "beta must be positive");
  if (d->sample_mode < DSM_SAMPLE_NONE || d->sample_mode > DSM_SAMPLE_MSD) return fail(DSM_ERR_ARG, "bad sample_mode");
  if ((d->sample_mode != DSM_SAMPLE_NONE) && !(d->delta > 0.0)) return fail(DSM_ERR_ARG, "delta must be positive");
  int rc = set_device(e);
  if (rc) return rc;
  e->gen++;
  Ens E;
  E.d = *d;
  E.d.nk_per_chain = nullptr;
  E.n = d->n_chains;
  E.nk.resize(E.n);
  int nmax = 0;
  for (int i = 0; i < E.n; ++i) {
    const int nk = d->nk_per_chain ? d->nk_per_chain[i] : d->nk;
    if (nk < 2 || nk > kMaxNk) return fail(DSM_ERR_ARG, "Nk=%d outside 2..%d", nk, kMaxNk);
    E.nk[i] = nk;
    nmax = std::max(nmax, nk);
    E.z_sum_est += ((double)nk + d->beta) / (1.0 + d->beta);
  }
  E.nmax = nmax;
  E.bin_max = bin_for(nmax);
  if (E.bin_max < 0) return fail(DSM_ERR_ARG, "Nk too large");
  // parameters
  const int nk_ref = d->nk_per_chain ? nmax : d->nk;
  Pcd pc;
  pcd_init(pc, d->beta, (float)nk_ref);
  // polydisperse CD background: partner quantiles and their sampler constants, shared by
  // ensembles with the same background (GEX parameters, or the same melt quantiles)
  float c_partner = 0.0f;
  if (d->partner_mode != DSM_PARTNER_OWN) {
    if (d->partner_mode != DSM_PARTNER_GEX && d->partner_mode != DSM_PARTNER_MELT)
      return fail(DSM_ERR_ARG, "bad partner_mode");
    dsm_engine::TabSlot key;
    key.kind = 1;
    key.mode = d->partner_mode;
    key.beta = d->beta;
    if (d->partner_mode == DSM_PARTNER_GEX) {
      std::copy(d->gex, d->gex + 4, key.gex);
    } else if (!partner_quantiles(DSM_PARTNER_MELT, d->gex, E.nk, key.nkF)) {
      return fail(DSM_ERR_ARG, "bad partner lengths");
    }
    bool ok = true;
    rc = take_tables(e, key, [&](std::vector<float>& data, dsm_engine::TabSlot& t) {
      std::vector<float> nkF = t.nkF;
      if (t.mode == DSM_PARTNER_GEX) ok = partner_quantiles(DSM_PARTNER_GEX, d->gex, E.nk, nkF);
      if (ok) t.c_create = partner_tables(nkF, d->beta, data);
    }, &E.pd_slot);
    if (!ok) return fail(DSM_ERR_ARG, "bad GEX parameters");
    if (rc) return rc;
    c_partner = e->tab_slots[E.pd_slot].c_create;
  }
  EnsParams& P = E.P;
  P.beta = d->beta;
  P.nk = nk_ref;
  P.nmax = nmax;
  P.cd = d->cd ? 1 : 0;
  P.flow = 0;
  for (int i = 0; i < 9; ++i) {
    P.kappa[i] = d->kappa[i];
    if (d->kappa[i] != 0.0f) P.flow = 1;
  }
  P.exp_wait = d->exp_wait ? 1 : 0;
  P.sample_mode = d->sample_mode;
  P.nchains = E.n;
  P.c_create = d->partner_mode != DSM_PARTNER_OWN ? c_partner : pc.normdt / d->beta;
  P.partner = d->partner_mode != DSM_PARTNER_OWN ? 1 : 0;
  P.pd_off = 0;
  P.s_At = pc.s_At; P.s_Ct = pc.s_Ct; P.s_Dt = pc.s_Dt; P.s_Adt = pc.s_Adt; P.s_Bdt = pc.s_Bdt;
  P.s_Cdt = pc.s_Cdt; P.s_Ddt = pc.s_Ddt; P.s_taud_inv = pc.s_taud_inv; P.s_one_minus_g = pc.s_one_minus_g;
  {
    const double pcut = 1.0 / (1.0 + (double)d->beta);
    double Tc = std::floor(pcut * 4294967296.0);
    if (Tc > 4294967295.0) Tc = 4294967295.0;
    P.cut_threshold = (uint32_t)Tc;
  }
  P.key0 = d->seed;
  P.key1 = d->salt;
  P.delta = d->delta;
  P.t_end = 0.0;
  P.k0 = 0;
  P.k_end = 0;
  // tables
  {
    dsm_engine::TabSlot key;
    key.kind = 0;
    key.nmax = nmax;
    key.cd = P.cd;
    key.beta = d->beta;
    key.c_create = P.c_create;
    rc = take_tables(e, key, [&](std::vector<float>& data, dsm_engine::TabSlot&) {
      build_tables(data, nmax, d->beta, P.cd, P.c_create);
    }, &E.tab_slot);
    if (rc) return rc;
  }
  P.tab_off = (int)e->tab_slots[E.tab_slot].off;
  P.fr_off = P.tab_off + kNumTables * (nmax + 2);
  if (P.partner) P.pd_off = (int)e->tab_slots[E.pd_slot].off;
  // sampling
  E.ncomp = (d->sample_mode == DSM_SAMPLE_GT || d->sample_mode == DSM_SAMPLE_MSD) ? 3
            : d->sample_mode == DSM_SAMPLE_FLOW ? 8 : 0;
  E.window = d->window > 0 ? d->window : e->opt.window;
  E.k_cur = d->sample_mode == DSM_SAMPLE_FLOW ? 1 : 0;
  E.p = d->corr_p > 0 ? d->corr_p : 64;
  E.m = d->corr_m > 0 ? d->corr_m : 8;
  if (E.p % E.m != 0 && eq_sampling(d->sample_mode)) return fail(DSM_ERR_ARG, "corr_p must be a multiple of corr_m");
  E.stage_floats = (int64_t)E.window * E.n * E.ncomp;
  P.stage_off = e->stage_space.take(E.stage_floats);
  e->stage_used = std::max(e->stage_used, e->stage_space.end);
  if ((rc = ensure_stage(e))) return rc;
  // chains
  E.chain0 = (int)e->chain_space.take(E.n);
  E.nstrand = 0;
  for (int i = 0; i < E.n; ++i) E.nstrand += cap_of(E.nk[i]);
  P.chain0 = E.chain0;
  E.strand0 = e->strand_space.take(E.nstrand);
  if ((rc = grow_chains(e, (int)e->chain_space.end, e->strand_space.end))) return rc;
  {
    const int id = (int)e->ens.size();
    std::vector<int> hnk(E.n), hens(E.n, id), hcap(E.n), hzero(E.n, 0), hloc(E.n);
    std::vector<uint32_t> hcid(E.n);
    std::vector<int64_t> hbase(E.n);
    int64_t b = E.strand0;
    for (int i = 0; i < E.n; ++i) {
      hnk[i] = E.nk[i]; hcap[i] = cap_of(E.nk[i]); hcid[i] = (uint32_t)i; hloc[i] = i; hbase[i] = b;
      b += cap_of(E.nk[i]);
    }
    const int c0 = E.chain0;
    HIPCHK(hipMemcpy(e->ch.nk + c0, hnk.data(), 4 * E.n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->ch.ens + c0, hens.data(), 4 * E.n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->ch.cap + c0, hcap.data(), 4 * E.n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->ch.cid + c0, hcid.data(), 4 * E.n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->ch.local + c0, hloc.data(), 4 * E.n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->ch.base + c0, hbase.data(), 8 * E.n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->ch.stall + c0, hzero.data(), 4 * E.n, hipMemcpyHostToDevice));
    HIPCHK(hipMemcpy(e->ch.Z + c0, hzero.data(), 4 * E.n, hipMemcpyHostToDevice));
    HIPCHK(hipMemset(e->ch.t + c0, 0, 8 * E.n));
    HIPCHK(hipMemset(e->ch.ev + c0, 0, 8 * E.n));
    HIPCHK(hipMemset(e->ch.zsum + c0, 0, 8 * E.n));
    HIPCHK(hipMemset(e->ch.R1 + 3 * (int64_t)c0, 0, 12 * (size_t)E.n));
    HIPCHK(hipMemset(e->ch.dt_prev + c0, 0, 4 * E.n));
    std::vector<int64_t> hns(E.n, E.k_cur);
    HIPCHK(hipMemcpy(e->ch.next_sample + c0, hns.data(), 8 * E.n, hipMemcpyHostToDevice));
    E.live = true;
    e->ens.push_back(std::move(E));
    if ((rc = upload_ens(e))) return rc;
    if (ens_id) *ens_id = id;
  }
  return DSM_OK;
}

int dsm_remove_ensemble(dsm_engine* e, int32_t ens) {
  Ens* E = get_ens(e, ens);
  if (!E) return fail(DSM_ERR_ARG, "no ensemble %d", ens);
  (void)set_device(e);
  e->gen++;
  if (E->dD) { (void)hipFree(E->dD); (void)hipFree(E->dA); (void)hipFree(E->dC); E->dD = E->dA = E->dC = nullptr; }
  E->live = false;
  E->flow_rows.clear();
  E->flow_rows.shrink_to_fit();
  // its chains, strand slots and staging rows go back to the engine's free ranges
  if (E->tab_slot >= 0) e->tab_slots[E->tab_slot].refs--;
  if (E->pd_slot >= 0) e->tab_slots[E->pd_slot].refs--;
  E->tab_slot = E->pd_slot = -1;
  e->chain_space.give(E->chain0, E->n);
  e->strand_space.give(E->strand0, E->nstrand);
  e->stage_space.give(E->P.stage_off, E->stage_floats);
  return DSM_OK;
}

int dsm_init_equilibrium(dsm_engine* e, int32_t ens) {
  Ens* E = get_ens(e, ens);
  if (!E) return fail(DSM_ERR_ARG, "no ensemble %d", ens);
  int rc = set_device(e);
  if (rc) return rc;
  HIPCHK(launch_iota(e->d_work, E->chain0, E->n, e->st));
  HIPCHK(launch_init_eq(e->ch, e->d_ens, e->d_tabs, e->d_work, E->n, e->st));
  HIPCHK(hipMemsetAsync(e->ch.R1 + 3 * (int64_t)E->chain0, 0, 12 * (size_t)E->n, e->st));  // common/chain.cu:207
  HIPCHK(hipStreamSynchronize(e->st));
  E->k_cur = E->P.sample_mode == kSampleFlow ? 1 : 0;
  E->flow_rows.clear();
  return DSM_OK;
}

int dsm_state_cap(dsm_engine* e, int32_t ens, int32_t* cap) {
  Ens* E = get_ens(e, ens);
  if (!E || !cap) return fail(DSM_ERR_ARG, "bad ensemble or pointer");
  *cap = E->nmax;
  return DSM_OK;
}

int dsm_ensemble_info(dsm_engine* e, int32_t ens, int32_t* n_chains, int32_t* cap, int32_t* sample_mode) {
  Ens* E = get_ens(e, ens);
  if (!E) return fail(DSM_ERR_ARG, "no ensemble %d", ens);
  if (n_chains) *n_chains = E->n;
  if (cap) *cap = E->nmax;
  if (sample_mode) *sample_mode = E->P.sample_mode;
  return DSM_OK;
}

int dsm_get_r1(dsm_engine* e, int32_t ens, float* r1) {
  Ens* E = get_ens(e, ens);
  if (!E || !r1) return fail(DSM_ERR_ARG, "bad ensemble or pointer");
  int rc = set_device(e);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(e->st));
  HIPCHK(hipMemcpy(r1, e->ch.R1 + 3 * (int64_t)E->chain0, 12 * (size_t)E->n, hipMemcpyDeviceToHost));
  return DSM_OK;
}

int dsm_get_state(dsm_engine* e, int32_t ens, int32_t cap, int32_t* Z, int32_t* N, float* Q, float* w, double* t,
                  uint64_t* ev, int32_t* stall, float* dt_prev, int64_t* next_sample) {
  Ens* E = get_ens(e, ens);
  if (!E) return fail(DSM_ERR_ARG, "no ensemble %d", ens);
  if (cap < E->nmax) return fail(DSM_ERR_ARG, "cap %d < Nk max %d", cap, E->nmax);
  int rc = set_device(e);
  if (rc) return rc;
  HIPCHK(hipStreamSynchronize(e->st));
  const int n = E->n, c0 = E->chain0;
  std::vector<int> hZ(n);
  HIPCHK(hipMemcpy(hZ.data(), e->ch.Z + c0, 4 * n, hipMemcpyDeviceToHost));
  if (Z) std::memcpy(Z, hZ.data(), 4 * n);
  if (t) HIPCHK(hipMemcpy(t, e->ch.t + c0, 8 * n, hipMemcpyDeviceToHost));
  if (ev) HIPCHK(hipMemcpy(ev, e->ch.ev + c0, 8 * n, hipMemcpyDeviceToHost));
  if (stall) HIPCHK(hipMemcpy(stall, e->ch.stall + c0, 4 * n, hipMemcpyDeviceToHost));
  if (dt_prev) HIPCHK(hipMemcpy(dt_prev, e->ch.dt_prev + c0, 4 * n, hipMemcpyDeviceToHost));
  if (next_sample) HIPCHK(hipMemcpy(next_sample, e->ch.next_sample + c0, 8 * n, hipMemcpyDeviceToHost));
  if (N || Q || w) {
    const int64_t ns = E->nstrand, s0 = E->strand0;
    std::vector<float> qx(ns), qy(ns), qz(ns), ww(ns);
    std::vector<int> nn(ns);
    HIPCHK(hipMemcpy(qx.data(), e->ch.Qx + s0, 4 * ns, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(qy.data(), e->ch.Qy + s0, 4 * ns, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(qz.data(), e->ch.Qz + s0, 4 * ns, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(ww.data(), e->ch.w + s0, 4 * ns, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(nn.data(), e->ch.N + s0, 4 * ns, hipMemcpyDeviceToHost));
    int64_t b = 0;
    for (int i = 0; i < n; ++i) {
      for (int k = 0; k < cap; ++k) {
        const bool on = k < hZ[i];
        const size_t o = (size_t)i * cap + k;
        if (N) N[o] = on ? nn[b + k] : 0;
        if (Q) {
          Q[o * 3 + 0] = on ? qx[b + k] : 0.0f;
          Q[o * 3 + 1] = on ? qy[b + k] : 0.0f;
          Q[o * 3 + 2] = on ? qz[b + k] : 0.0f;
        }
        if (w) w[o] = (k < hZ[i] - 1) ? ww[b + k] : 0.0f;
      }
      b += cap_of(E->nk[i]);
    }
  }
  return DSM_OK;
}

int dsm_set_state(dsm_engine* e, int32_t ens, int32_t cap, const int32_t* Z, const int32_t* N, const float* Q,
                  const float* w, const double* t, const uint64_t* ev, const int32_t* stall, const float* dt_prev,
                  const int64_t* next_sample) {
  Ens* E = get_ens(e, ens);
  if (!E || !Z || !N || !Q || !w) return fail(DSM_ERR_ARG, "bad ensemble or null state array");
  if (cap < E->nmax) return fail(DSM_ERR_ARG, "cap %d < Nk max %d", cap, E->nmax);
  const int n = E->n, c0 = E->chain0;
  for (int i = 0; i < n; ++i) {
    if (Z[i] < 1 || Z[i] > E->nk[i]) return fail(DSM_ERR_ARG, "chain %d: Z=%d outside 1..%d", i, Z[i], E->nk[i]);
    int64_t sum = 0;
    for (int k = 0; k < Z[i]; ++k) {
      const int v = N[(size_t)i * cap + k];
      if (v < 1) return fail(DSM_ERR_ARG, "chain %d strand %d: N=%d < 1", i, k, v);
      sum += v;
    }
    if (sum != E->nk[i]) return fail(DSM_ERR_ARG, "chain %d: sum N = %lld != Nk %d", i, (long long)sum, E->nk[i]);
  }
  int rc = set_device(e);
  if (rc) return rc;
  const int64_t ns = E->nstrand, s0 = E->strand0;
  std::vector<float> qx(ns, 0.0f), qy(ns, 0.0f), qz(ns, 0.0f), ww(ns, 0.0f);
  std::vector<int> nn(ns, 0);
  int64_t b = 0;
  for (int i = 0; i < n; ++i) {
    for (int k = 0; k < Z[i]; ++k) {
      const size_t o = (size_t)i * cap + k;
      nn[b + k] = N[o]; qx[b + k] = Q[o * 3 + 0]; qy[b + k] = Q[o * 3 + 1]; qz[b + k] = Q[o * 3 + 2];
      ww[b + k] = (k < Z[i] - 1) ? w[o] : 0.0f;
    }
    b += cap_of(E->nk[i]);
  }
  HIPCHK(hipStreamSynchronize(e->st));
  HIPCHK(hipMemcpy(e->ch.Qx + s0, qx.data(), 4 * ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->ch.Qy + s0, qy.data(), 4 * ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->ch.Qz + s0, qz.data(), 4 * ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->ch.w + s0, ww.data(), 4 * ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->ch.N + s0, nn.data(), 4 * ns, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->ch.Z + c0, Z, 4 * n, hipMemcpyHostToDevice));
  std::vector<double> zt(n, 0.0);
  std::vector<uint64_t> zev(n, 0);
  std::vector<int> zst(n, 0);
  std::vector<float> zdt(n, 0.0f);
  std::vector<int64_t> zns(n, E->k_cur);
  HIPCHK(hipMemcpy(e->ch.t + c0, t ? t : zt.data(), 8 * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->ch.ev + c0, ev ? ev : zev.data(), 8 * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->ch.stall + c0, stall ? stall : zst.data(), 4 * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->ch.dt_prev + c0, dt_prev ? dt_prev : zdt.data(), 4 * n, hipMemcpyHostToDevice));
  HIPCHK(hipMemcpy(e->ch.next_sample + c0, next_sample ? next_sample : zns.data(), 8 * n, hipMemcpyHostToDevice));
  if (next_sample) E->k_cur = next_sample[0];
  return DSM_OK;
}

int dsm_advance(dsm_engine* e, int32_t ens, double t_end, int32_t mode, volatile const int32_t* run_flag,
                volatile int32_t* progress, dsm_run_stats* stats) {
  if (!get_ens(e, ens)) return fail(DSM_ERR_ARG, "no ensemble %d", ens);
  if (mode != DSM_MODE_BATCHED && mode != DSM_MODE_PLAIN) return fail(DSM_ERR_ARG, "bad mode %d", mode);
  return advance_impl(e, {ens}, {t_end}, mode, run_flag, progress, stats);
}

int dsm_advance_many(dsm_engine* e, const int32_t* ens_ids, const double* t_end, int32_t count, int32_t mode,
                     volatile const int32_t* run_flag, dsm_run_stats* stats) {
  if (!e || !ens_ids || !t_end || count <= 0) return fail(DSM_ERR_ARG, "bad arguments");
  std::vector<int> ids(ens_ids, ens_ids + count);
  std::vector<double> te(t_end, t_end + count);
  for (int id : ids)
    if (!get_ens(e, id)) return fail(DSM_ERR_ARG, "no ensemble %d", id);
  return advance_impl(e, ids, te, mode, run_flag, nullptr, stats);
}

int dsm_advance_events(dsm_engine* e, int32_t ens, int64_t n_events, int32_t mode, dsm_run_stats* stats) {
  Ens* E = get_ens(e, ens);
  if (!E) return fail(DSM_ERR_ARG, "no ensemble %d", ens);
  if (n_events < 0) return fail(DSM_ERR_ARG, "negative event count");
  int rc = set_device(e);
  if (rc) return rc;
  Timer T;
  dsm_run_stats rs{};
  unsigned long long before[2];
  if ((rc = chain_totals(e, {ens}, before))) return rc;
  E->P.t_end = INFINITY;
  E->P.k0 = E->k_cur;
  E->P.k_end = E->k_cur;  // sampling suspended
  if ((rc = upload_ens(e))) return rc;
  if (mode == DSM_MODE_PLAIN) {
    // every chain stops at its jump target; each launch makes up to plain_jumps of them
    HIPCHK(launch_set_ev_stop(e->ch.ev, e->ch.ev_stop, E->chain0, E->n, (uint64_t)n_events, e->st));
    Pass ps;
    if ((rc = pass_setup(e, {ens}, true, ps))) return rc;
    const int64_t launches = (n_events + ps.a.plain_jumps - 1) / ps.a.plain_jumps;
    HIPCHK(hipEventRecord(e->evs[2], e->st));
    for (int64_t k = 0; k < launches; ++k)
      if ((rc = pass_bins(e, ps, true, false, &rs))) return rc;
    HIPCHK(hipEventRecord(e->evs[3], e->st));
    HIPCHK(hipEventSynchronize(e->evs[3]));
    float ms = 0.0f;
    HIPCHK(hipEventElapsedTime(&ms, e->evs[2], e->evs[3]));
    rs.kmc_ms = ms;
    rs.first_kmc_ms = ms;
  } else {
    HIPCHK(launch_set_ev_stop(e->ch.ev, e->ch.ev_stop, E->chain0, E->n, (uint64_t)n_events, e->st));
    if ((rc = kmc_pass(e, {ens}, mode, true, &rs))) return rc;
  }
  if ((rc = finish_stats(e, {ens}, before, &rs, T))) return rc;
  if (stats) *stats = rs;
  return DSM_OK;
}

int dsm_flow_stress(dsm_engine* e, int32_t ens, int64_t first_row, int64_t max_rows, double* out, int64_t* n_rows) {
  Ens* E = get_ens(e, ens);
  if (!E) return fail(DSM_ERR_ARG, "no ensemble %d", ens);
  const int64_t have = (int64_t)E->flow_rows.size() / 9;
  if (n_rows) *n_rows = have;
  if (out && first_row < have) {
    const int64_t k = std::min(max_rows, have - first_row);
    std::memcpy(out, E->flow_rows.data() + first_row * 9, sizeof(double) * 9 * k);
  }
  return DSM_OK;
}

int dsm_stress(dsm_engine* e, int32_t ens, double* out, int32_t* healthy) {
  Ens* E = get_ens(e, ens);
  if (!E || !out) return fail(DSM_ERR_ARG, "bad ensemble or pointer");
  int rc = set_device(e);
  if (rc) return rc;
  HIPCHK(launch_chain_stress(e->ch, E->chain0, E->n, E->P.kappa, e->d_part, e->st));
  std::vector<double> part((size_t)kStressBlocks * 9);
  HIPCHK(hipMemcpyAsync(part.data(), e->d_part, sizeof(double) * part.size(), hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  double sum[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < kStressBlocks; ++b)
    for (int q = 0; q < 9; ++q) sum[q] += part[(size_t)b * 9 + q];
  for (int q = 0; q < 8; ++q) out[q] = sum[8] > 0.0 ? sum[q] / sum[8] : 0.0;
  if (healthy) *healthy = (int32_t)sum[8];
  return DSM_OK;
}

int dsm_set_option(dsm_engine* e, int32_t key, int32_t value) {
  if (!e) return fail(DSM_ERR_ARG, "null engine");
  EngineOptions& o = e->opt;
  e->gen++;  // captured passes hold the options they were captured with
  switch (key) {
    case DSM_OPT_PLAIN_JUMPS: if (value < 1) break; o.plain_jumps = value; return DSM_OK;
    case DSM_OPT_SMALL_PER_CU: if (value < 0) break; o.small_per_cu = value; return DSM_OK;
    case DSM_OPT_SPLIT_MARGIN: if (value < 0) break; o.split_margin_div = value; return DSM_OK;
    case DSM_OPT_LDS_TABLE_FLOATS: if (value < 0) break; o.lds_table_floats = value; return DSM_OK;
    case DSM_OPT_CONCURRENT: if (value < -1 || value > 1) break; o.concurrent = value; return DSM_OK;
    case DSM_OPT_BLOCKS_PER_CU: if (value < 0) break; o.blocks_per_cu = value; return DSM_OK;
    case DSM_OPT_WINDOW: if (value < 1) break; o.window = value; return DSM_OK;
    case DSM_OPT_GRAPH: if (value < 0 || value > 1) break; o.graph = value; return DSM_OK;
    case DSM_OPT_WIDE_PAIR: if (value < 0 || value > 1) break; o.wide_pair = value; return DSM_OK;
    default: return fail(DSM_ERR_ARG, "unknown option %d", key);
  }
  return fail(DSM_ERR_ARG, "option %d: value %d out of range", key, value);
}

int dsm_get_option(dsm_engine* e, int32_t key, int32_t* value) {
  if (!e || !value) return fail(DSM_ERR_ARG, "null argument");
  const EngineOptions& o = e->opt;
  switch (key) {
    case DSM_OPT_PLAIN_JUMPS: *value = o.plain_jumps; return DSM_OK;
    case DSM_OPT_SMALL_PER_CU: *value = o.small_per_cu; return DSM_OK;
    case DSM_OPT_SPLIT_MARGIN: *value = o.split_margin_div; return DSM_OK;
    case DSM_OPT_LDS_TABLE_FLOATS: *value = o.lds_table_floats; return DSM_OK;
    case DSM_OPT_CONCURRENT: *value = o.concurrent; return DSM_OK;
    case DSM_OPT_BLOCKS_PER_CU: *value = o.blocks_per_cu; return DSM_OK;
    case DSM_OPT_WINDOW: *value = o.window; return DSM_OK;
    case DSM_OPT_GRAPH: *value = o.graph; return DSM_OK;
    case DSM_OPT_WIDE_PAIR: *value = o.wide_pair; return DSM_OK;
    default: return fail(DSM_ERR_ARG, "unknown option %d", key);
  }
}

int dsm_check_state(dsm_engine* e, int32_t ens, double t_min, int64_t* out) {
  Ens* E = get_ens(e, ens);
  if (!E || !out) return fail(DSM_ERR_ARG, "bad ensemble or pointer");
  int rc = set_device(e);
  if (rc) return rc;
  unsigned long long* cnt = (unsigned long long*)e->d_part;  // the stress partials' space, 7 counters
  HIPCHK(launch_fill_u32(cnt, 0u, 14, e->st));
  HIPCHK(launch_check_state(e->ch, E->chain0, E->n, t_min, cnt, e->st));
  unsigned long long h[7];
  HIPCHK(hipMemcpyAsync(h, cnt, sizeof(h), hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  for (int q = 0; q < 7; ++q) out[q] = (int64_t)h[q];
  return DSM_OK;
}

int dsm_correlator(dsm_engine* e, int32_t ens, int32_t max_pts, double* lag, double* g, int32_t* npts) {
  Ens* E = get_ens(e, ens);
  if (!E) return fail(DSM_ERR_ARG, "no ensemble %d", ens);
  return dsm_correlator_range(e, ens, 0, E->n, max_pts, lag, g, npts);
}

int dsm_correlator_range(dsm_engine* e, int32_t ens, int32_t first, int32_t count, int32_t max_pts, double* lag,
                         double* g, int32_t* npts) {
  Ens* E = get_ens(e, ens);
  if (!E) return fail(DSM_ERR_ARG, "no ensemble %d", ens);
  if (!eq_sampling(E->P.sample_mode) || !E->dC)
    return fail(DSM_ERR_STATE, "ensemble %d has no G(t) or MSD(t) correlator", ens);
  if (first < 0 || count < 1 || first > E->n - count)
    return fail(DSM_ERR_ARG, "chain range %d+%d outside 0..%d", first, count, E->n);
  int rc = set_device(e);
  if (rc) return rc;
  const int s = E->s, p = E->p, m = E->m;
  HIPCHK(launch_corr_reduce(E->dC, E->n, e->ch.stall, E->chain0, first, count, s * p, e->d_red, e->d_healthy, e->st));
  std::vector<double> sums(s * p);
  int healthy = 0;
  HIPCHK(hipMemcpyAsync(sums.data(), e->d_red, sizeof(double) * s * p, hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipMemcpyAsync(&healthy, e->d_healthy, sizeof(int), hipMemcpyDeviceToHost, e->st));
  HIPCHK(hipStreamSynchronize(e->st));
  int im = 0;
  double mk = 1.0;
  for (int k = 0; k < s; ++k) {
    const int jmin = k == 0 ? 0 : p / m;
    for (int j = jmin; j < p; ++j) {
      const double npair = (double)std::max<int64_t>(0, E->nins[k] - j);
      if (im < max_pts) {
        if (lag) lag[im] = (double)j * mk;
        // G: trace / 3 (common/correlator.cu:176); MSD: sum of the components (:177)
        const double norm = E->P.sample_mode == kSampleMsd ? 1.0 : 3.0;
        if (g) g[im] = (npair > 0 && healthy > 0) ? sums[k * p + j] / npair / norm / healthy : 0.0;
      }
      ++im;
    }
    mk *= m;
  }
  if (npts) *npts = im;
  return DSM_OK;
}

int dsm_last_samples(dsm_engine* e, int32_t ens, float* out, int64_t max_floats, int64_t* k0, int64_t* nrows) {
  Ens* E = get_ens(e, ens);
  if (!E) return fail(DSM_ERR_ARG, "no ensemble %d", ens);
  const int64_t nf = E->last_rows * E->n * E->ncomp;
  if (k0) *k0 = E->last_k0;
  if (nrows) *nrows = E->last_rows;
  if (out && nf > 0) {
    HIPCHK(hipSetDevice(e->device));
    HIPCHK(hipStreamSynchronize(e->st));
    HIPCHK(hipMemcpy(out, e->d_stage + E->P.stage_off, sizeof(float) * std::min(nf, max_floats),
                     hipMemcpyDeviceToHost));
  }
  return DSM_OK;
}

int dsm_tables(dsm_engine* e, int32_t ens, int32_t which, float* out, int32_t* n) {
  Ens* E = get_ens(e, ens);
  if (!E || which < 0 || which > 13) return fail(DSM_ERR_ARG, "bad ensemble or table");
  if (which >= 10 && !E->P.partner) return fail(DSM_ERR_STATE, "ensemble %d has no partner tables", ens);
  const int nn = E->nmax + 2;
  // 10/11: f_t / f_d_t partner sampler constants (5 per quantile); 12/13: their Nk
  const int cnt = which == 9 ? 4 * E->nmax + 4 : which >= 12 ? kPartnerK : which >= 10 ? 5 * kPartnerK : nn;
  const int off = which == 9    ? E->P.fr_off
                  : which >= 12 ? E->P.pd_off + 10 * kPartnerK + (which - 12) * kPartnerK
                  : which >= 10 ? E->P.pd_off + (which - 10) * 5 * kPartnerK
                                : E->P.tab_off + which * nn;
  if (n) *n = cnt;
  if (out) std::memcpy(out, e->tabs.data() + off, sizeof(float) * cnt);
  return DSM_OK;
}

int dsm_ensemble_scalars(dsm_engine* e, int32_t ens, float* o) {
  Ens* E = get_ens(e, ens);
  if (!E || !o) return fail(DSM_ERR_ARG, "bad ensemble or pointer");
  Pcd pc;
  pcd_init(pc, E->d.beta, (float)E->P.nk);
  const float v[24] = {pc.g, pc.alpha, pc.tau_0, pc.tau_max, pc.tau_d, pc.c1, pc.cm1, pc.rcm1c1,
                       pc.At, pc.Adt, pc.Bdt, pc.normdt, pc.s_At, pc.s_Ct, pc.s_Dt, pc.s_Adt,
                       pc.s_Bdt, pc.s_Cdt, pc.s_Ddt, pc.s_taud_inv, pc.s_one_minus_g, E->P.c_create,
                       (float)E->P.cut_threshold, E->P.beta};
  std::memcpy(o, v, sizeof(v));
  return DSM_OK;
}

// ------------------------------------------------------------------ probes
static int probe_run(int32_t device_id, const void* in, size_t in_bytes, void* out, size_t out_bytes,
                     hipError_t (*launch)(const void*, void*, int, int, hipStream_t), int n, int which) {
  HIPCHK(hipSetDevice(device_id));
  void *din = nullptr, *dout = nullptr;
  HIPCHK(hipMalloc(&din, in_bytes));
  HIPCHK(hipMalloc(&dout, out_bytes));
  HIPCHK(hipMemcpy(din, in, in_bytes, hipMemcpyHostToDevice));
  HIPCHK(hipMemset(dout, 0, out_bytes));
  HIPCHK(launch(din, dout, n, which, nullptr));
  HIPCHK(hipDeviceSynchronize());
  HIPCHK(hipMemcpy(out, dout, out_bytes, hipMemcpyDeviceToHost));
  HIPCHK(hipFree(din));
  HIPCHK(hipFree(dout));
  return DSM_OK;
}

static hipError_t l_scan(const void* i, void* o, int n, int form, hipStream_t s) {
  return launch_scan_probe((const float*)i, (float*)o, n, form, s);
}
static hipError_t l_com(const void* i, void* o, int n, int g, hipStream_t s) {
  return launch_com_probe((const float*)i, (float*)o, n, g, s);
}
static hipError_t l_math(const void* i, void* o, int n, int w, hipStream_t s) {
  return launch_math_probe((const float*)i, (float*)o, n, w, s);
}
static hipError_t l_philox(const void* i, void* o, int n, int, hipStream_t s) {
  return launch_philox_probe((const uint32_t*)i, (uint32_t*)o, n, s);
}

int dsm_probe_scan(int32_t device_id, const float* in, float* out, int32_t n, int32_t form) {
  if (!in || !out || form < 0 || form > 5 || n <= 0 ||
      n > (form == 1 ? 1024 : form == 2 ? 128 : form == 3 ? 256 : form == 4 ? 128 : form == 5 ? 512 : 64))
    return fail(DSM_ERR_ARG, "bad probe arguments");
  const int chains = (form == 1 || form == 5) ? 1 : (form == 2 || form == 3) ? 2 : 4;
  const int nin = chains * n, nout = chains * (n + 1);
  return probe_run(device_id, in, 4 * (size_t)nin, out, 4 * (size_t)nout, l_scan, n, form);
}
int dsm_probe_com(int32_t device_id, const float* in, float* out, int32_t n, int32_t geom) {
  if (!in || !out || geom < 0 || (geom > 4 && geom != kGeomPair2 && geom != kGeomRow2) || n < 1 ||
      n > kmc_geom_capacity(geom))
    return fail(DSM_ERR_ARG, "bad probe arguments");
  return probe_run(device_id, in, 4 * (size_t)(4 * n + 1), out, 4 * 3, l_com, n, geom);
}
int dsm_probe_math(int32_t device_id, const float* in, float* out, int32_t n, int32_t which) {
  if (n <= 0 || !in || !out || which < 0 || which > 3) return fail(DSM_ERR_ARG, "bad probe arguments");
  return probe_run(device_id, in, 4 * (size_t)n, out, 4 * (size_t)n, l_math, n, which);
}
int dsm_probe_philox(int32_t device_id, const uint32_t* ck, uint32_t* out, int32_t n) {
  if (n <= 0 || !ck || !out) return fail(DSM_ERR_ARG, "bad probe arguments");
  return probe_run(device_id, ck, 24 * (size_t)n, out, 16 * (size_t)n, l_philox, n, 0);
}

int dsm_probe_copy(int32_t device_id, int64_t n_bytes, int32_t reps, double* ms_per_copy) {
  if (n_bytes < 16 || (n_bytes & 15) || reps <= 0 || !ms_per_copy) return fail(DSM_ERR_ARG, "bad probe arguments");
  HIPCHK(hipSetDevice(device_id));
  hipDeviceProp_t prop;
  HIPCHK(hipGetDeviceProperties(&prop, device_id));
  void *a = nullptr, *b = nullptr;
  HIPCHK(hipMalloc(&a, (size_t)n_bytes));
  if (hipMalloc(&b, (size_t)n_bytes) != hipSuccess) {
    (void)hipFree(a);
    return fail(DSM_ERR_MEMORY, "copy probe: out of device memory");
  }
  const size_t n4 = (size_t)n_bytes / 16;
  const int grid = prop.multiProcessorCount * 16;
  hipEvent_t e0, e1;
  HIPCHK(hipEventCreate(&e0));
  HIPCHK(hipEventCreate(&e1));
  HIPCHK(hipMemset(a, 0, (size_t)n_bytes));
  HIPCHK(launch_copy_probe(a, b, n4, grid, nullptr));  // warm-up
  HIPCHK(hipEventRecord(e0, nullptr));
  for (int r = 0; r < reps; ++r) HIPCHK(launch_copy_probe(a, b, n4, grid, nullptr));
  HIPCHK(hipEventRecord(e1, nullptr));
  HIPCHK(hipEventSynchronize(e1));
  float ms = 0.0f;
  HIPCHK(hipEventElapsedTime(&ms, e0, e1));
  *ms_per_copy = (double)ms / reps;
  HIPCHK(hipEventDestroy(e0));
  HIPCHK(hipEventDestroy(e1));
  HIPCHK(hipFree(a));
  HIPCHK(hipFree(b));
  return DSM_OK;
}

}  // extern "C"
