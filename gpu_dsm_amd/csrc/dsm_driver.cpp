// dsm_driver.cpp -- drop-in replacement for main_cuda (ktaletsk/gpu_dsm
// common/main_cuda.cu:34-185): input.dat parsing, flow / G(t) dispatch, the
// tau/G[_ID].dat, distr_*.dat and -s/-l binary formats, stdout lines.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <fstream>
#include <iostream>
#include <sstream>
#include <string>
#include <vector>

#include "dsm.h"

namespace {

// common/job_ID.h:33-68: "<name>_<id>.dat", or "<name>.dat" for job_ID 0
std::string filename_ID(const std::string& dir, const std::string& name, int job_ID) {
  std::string base = dir.empty() ? std::string() : dir + "/";
  if (job_ID != 0) return base + name + "_" + std::to_string(job_ID) + ".dat";
  return base + name + ".dat";
}

struct Input {
  float Be = 1.0f;
  int NK = 0;
  int N_cha = 0;
  float kappa[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  int CD_flag = 0;
  int PD_flag = 0;
  int equilibrium_type = 0;
  double t_step_size = 1.0;
  double simulation_time = 0.0;
};

// common/main_cuda.cu:53-68: nine whitespace-separated values in order
bool read_input(const std::string& path, Input& in) {
  std::ifstream f(path);
  if (!f.is_open()) return false;
  f >> in.Be >> in.NK >> in.N_cha;
  for (float& k : in.kappa) f >> k;
  f >> in.CD_flag >> in.PD_flag >> in.equilibrium_type >> in.t_step_size >> in.simulation_time;
  return !f.fail();
}

// common/chain.h:45-53
struct ChainHead {
  int32_t Z;
  float time;
  float dummy;
  int32_t stall_flag;
};

struct HostState {
  int n = 0, cap = 0;
  std::vector<int32_t> Z, N, stall;
  std::vector<float> Q, w, dt;
  std::vector<double> t;
  std::vector<uint64_t> ev;
  std::vector<int64_t> ns;
  void resize(int nn, int cc) {
    n = nn; cap = cc;
    Z.assign(n, 0); N.assign((size_t)n * cap, 0); stall.assign(n, 0);
    Q.assign((size_t)n * cap * 3, 0.0f); w.assign((size_t)n * cap, 0.0f); dt.assign(n, 0.0f);
    t.assign(n, 0.0); ev.assign(n, 0); ns.assign(n, 0);
  }
};

int fetch(dsm_engine* e, int ens, HostState& h) {
  int cap = 0;
  int rc = dsm_state_cap(e, ens, &cap);
  if (rc) return rc;
  return dsm_get_state(e, ens, h.cap, h.Z.data(), h.N.data(), h.Q.data(), h.w.data(), h.t.data(), h.ev.data(),
                       h.stall.data(), h.dt.data(), h.ns.data());
}

thread_local std::string g_io_err;

// common/ensemble.cu:342-354, common/chain.cu:236-242
bool save_conformations(const std::string& path, const HostState& h, double t_ref) {
  std::ofstream f(path, std::ios::out | std::ios::binary);
  if (!f.is_open()) return false;
  const int32_t n = h.n;
  f.write((const char*)&n, 4);
  for (int i = 0; i < h.n; ++i) {
    ChainHead hd{h.Z[i], (float)(h.t[i] - t_ref), 0.0f, h.stall[i]};
    f.write((const char*)&hd, sizeof(hd));
    for (int k = 0; k < h.Z[i]; ++k) {
      const size_t o = (size_t)i * h.cap + k;
      const float q[4] = {h.Q[o * 3], h.Q[o * 3 + 1], h.Q[o * 3 + 2], (float)h.N[o]};
      f.write((const char*)q, 16);
    }
    for (int k = 0; k < h.Z[i]; ++k) {
      const float r = (k < h.Z[i] - 1) ? h.w[(size_t)i * h.cap + k] : 0.0f;
      f.write((const char*)&r, 4);
    }
  }
  return f.good();
}

// common/ensemble.cu:486-504, common/chain.cu:244-250.  The head's unused float and the
// rate slot of the last strand are read and ignored (the reference leaves them unset).
int load_conformations(const std::string& path, HostState& h, int n_expected) {
  std::ifstream f(path, std::ios::in | std::ios::binary);
  if (!f.is_open()) { g_io_err = "cannot open " + path; return DSM_ERR_IO; }
  int32_t n = 0;
  f.read((char*)&n, 4);
  if (!f) { g_io_err = path + ": no chain count"; return DSM_ERR_IO; }
  if (n != n_expected) {
    g_io_err = path + ": ensemble size mismatch (" + std::to_string(n) + " chains, " + std::to_string(n_expected) +
               " expected)";
    return DSM_ERR_MISMATCH;
  }
  for (int i = 0; i < h.n; ++i) {
    ChainHead hd;
    f.read((char*)&hd, sizeof(hd));
    if (!f || hd.Z < 1 || hd.Z > h.cap) {
      g_io_err = path + ": chain " + std::to_string(i) + (f ? ": Z out of range" : ": file ends inside its record");
      return DSM_ERR_IO;
    }
    h.Z[i] = hd.Z;
    h.t[i] = (double)hd.time;
    h.stall[i] = hd.stall_flag;
    for (int k = 0; k < hd.Z; ++k) {
      float q[4];
      f.read((char*)q, 16);
      const size_t o = (size_t)i * h.cap + k;
      h.Q[o * 3] = q[0]; h.Q[o * 3 + 1] = q[1]; h.Q[o * 3 + 2] = q[2];
      h.N[o] = (int32_t)q[3];
    }
    for (int k = 0; k < hd.Z; ++k) {
      float r;
      f.read((char*)&r, 4);
      h.w[(size_t)i * h.cap + k] = (k < hd.Z - 1) ? r : 0.0f;
    }
    if (!f) { g_io_err = path + ": chain " + std::to_string(i) + ": file ends inside its record"; return DSM_ERR_IO; }
  }
  return DSM_OK;
}

// common/ensemble.cu:356-392 (Z pdf; P sized by Zmax rather than N_cha)
void save_Z_distribution(const std::string& path, const HostState& h) {
  std::ofstream f(path);
  if (!f.is_open() || h.n == 0) return;
  int Zmin = h.Z[0], Zmax = h.Z[0];
  for (int z : h.Z) { Zmin = std::min(Zmin, z); Zmax = std::max(Zmax, z); }
  std::vector<int> cnt(Zmax + 1, 0);
  for (int z : h.Z) cnt[z]++;
  for (int j = Zmin; j <= Zmax; ++j) f << j << "\t" << (float)cnt[j] / h.n << "\n";
}

// common/ensemble.cu:394-430 (N pdf over all strands)
void save_N_distribution(const std::string& path, const HostState& h) {
  std::ofstream f(path);
  if (!f.is_open() || h.n == 0) return;
  int Nmin = h.N[0], Nmax = h.N[0];
  long Nstr = 0;
  for (int i = 0; i < h.n; ++i) {
    for (int k = 0; k < h.Z[i]; ++k) {
      const int v = h.N[(size_t)i * h.cap + k];
      Nmin = std::min(Nmin, v);
      Nmax = std::max(Nmax, v);
      if (v == 0) std::cout << "Zero length strand in chain " << i << ", #" << k << "\n";
    }
    Nstr += h.Z[i];
  }
  std::vector<long> cnt(Nmax + 1, 0);
  for (int i = 0; i < h.n; ++i)
    for (int k = 0; k < h.Z[i]; ++k) cnt[h.N[(size_t)i * h.cap + k]]++;
  for (int v = Nmin; v <= Nmax; ++v) {
    const float P = (float)cnt[v] / Nstr;
    f << v << "\t" << P << "\n";
  }
}

// common/ensemble.cu:438-484 (CDF of |Q| over internal strands, every 10th point)
void save_Q_distribution(const std::string& path, const HostState& h) {
  std::ofstream f(path);
  if (!f.is_open()) return;
  std::vector<float> Q;
  for (int i = 0; i < h.n; ++i)
    for (int k = 1; k < h.Z[i] - 1; ++k) {
      const size_t o = (size_t)i * h.cap + k;
      const float x = h.Q[o * 3], y = h.Q[o * 3 + 1], z = h.Q[o * 3 + 2];
      Q.push_back(std::sqrt(x * x + y * y + z * z));
    }
  std::cout << "\n";
  std::sort(Q.begin(), Q.end());
  const size_t Nstr = Q.size();
  std::vector<float> P(Nstr);
  for (size_t i = 0; i < Nstr; ++i) {
    if (i != 0 && Q[i] == Q[i - 1]) P[i - 1] = (float)i / (float)Nstr;
    P[i] = (float)i / (float)Nstr;
  }
  const size_t quant = 10;
  for (size_t i = 0; i < Nstr / quant; ++i) f << Q[i * quant] << "\t" << P[i * quant] << "\n";
}

}  // namespace

extern "C" int dsm_save(dsm_engine* e, int32_t ens, const char* path, double t_ref) {
  int32_t n = 0, cap = 0;
  int rc = dsm_ensemble_info(e, ens, &n, &cap, nullptr);
  if (rc) return rc;
  if (!path) return DSM_ERR_ARG;
  HostState h;
  h.resize(n, cap);
  if ((rc = fetch(e, ens, h))) return rc;
  return save_conformations(path, h, t_ref) ? DSM_OK : DSM_ERR_IO;
}

extern "C" int dsm_load(dsm_engine* e, int32_t ens, const char* path) {
  int32_t n = 0, cap = 0, mode = 0;
  int rc = dsm_ensemble_info(e, ens, &n, &cap, &mode);
  if (rc) return rc;
  if (!path) return DSM_ERR_ARG;
  HostState h;
  h.resize(n, cap);
  g_io_err.clear();
  if ((rc = load_conformations(path, h, n))) return rc;
  // sampling restarts at the loaded clocks: flow output 1 at delta, G(t) sample 0
  for (int i = 0; i < n; ++i) h.ns[i] = mode == DSM_SAMPLE_FLOW ? 1 : 0;
  rc = dsm_set_state(e, ens, cap, h.Z.data(), h.N.data(), h.Q.data(), h.w.data(), h.t.data(), nullptr,
                     h.stall.data(), nullptr, h.ns.data());
  if (rc == DSM_ERR_ARG) {
    g_io_err = std::string(path) + ": invalid chain record (" + dsm_last_error() + ")";
    return DSM_ERR_IO;
  }
  return rc;
}

extern "C" const char* dsm_load_error(void) { return g_io_err.c_str(); }

extern "C" int dsm_main_ex(volatile const bool* run_flag, int job_ID, const char* savefile, const char* loadfile,
                           int device_ID, bool distr, volatile int* progress_bar, const dsm_main_options* opt) {
  dsm_main_options o{};
  o.mode = DSM_MODE_BATCHED;
  o.corr_p = 64;
  o.corr_m = 8;
  if (opt) {
    o = *opt;
    if (o.corr_p <= 0) o.corr_p = 64;
    if (o.corr_m <= 0) o.corr_m = 8;
  }
  const std::string dir = o.workdir ? o.workdir : "";
  if (job_ID != 0) {
    std::cout << "job_ID: " << job_ID << '\n';
    std::cout << "using " << job_ID << " as a seed for random number generator\n";
    std::cout << "\"_" << job_ID << "\" will be appended to filename for all files generated\n\n";
  }
  Input in;
  if (!read_input(dir.empty() ? "input.dat" : dir + "/input.dat", in)) {
    std::cout << "cannot read input.dat\n";
    return DSM_ERR_IO;
  }
  const int s = (int)std::ceil(std::log(in.simulation_time / in.t_step_size / o.corr_p) / std::log((double)o.corr_m)) + 1;
  std::cout << "\nsimulation parameters:\n";
  std::cout << "NK\tBeta\tN_cha" << "\n";
  std::cout << in.NK << '\t' << in.Be << '\t' << in.N_cha << "\n";
  std::cout << "deformation tensor:" << "\n";
  std::cout << in.kappa[0] << '\t' << in.kappa[1] << '\t' << in.kappa[2] << '\n' << in.kappa[3] << '\t' << in.kappa[4]
            << '\t' << in.kappa[5] << '\n' << in.kappa[6] << '\t' << in.kappa[7] << '\t' << in.kappa[8] << '\n';
  if (in.equilibrium_type == 1) std::cout << "G(t) calculation is on\n";
  else if (in.equilibrium_type == 2) std::cout << "MSD(t) calculation is on\n";
  std::cout << "simulation time, sync time" << "\n";
  std::cout << in.simulation_time << '\t' << in.t_step_size << '\n';
  std::cout << "number of correlator levels" << '\t' << s << '\n' << '\n';

  bool flow = false;
  for (float k : in.kappa) flow = flow || (k != 0.0f);
  // common/main_cuda.cu:95-103: PD_flag reads the GEX parameters a, b, mp, Mk
  float gex[4] = {0.0f, 0.0f, 0.0f, 0.0f};
  if (in.PD_flag) {
    std::ifstream pf(dir.empty() ? "polydisp.dat" : dir + "/polydisp.dat");
    if (!(pf >> gex[0] >> gex[1] >> gex[2] >> gex[3])) {
      std::cout << "cannot read polydisp.dat\n";
      return DSM_ERR_IO;
    }
  }
  dsm_engine* e = nullptr;
  int rc = dsm_create(device_ID, &e);
  if (rc) {
    std::cout << "device error: " << dsm_last_error() << "\n";
    return rc;
  }
  dsm_ensemble_desc d{};
  d.n_chains = in.N_cha;
  d.nk = in.NK;
  d.beta = in.Be;
  d.cd = in.CD_flag;
  for (int i = 0; i < 9; ++i) d.kappa[i] = in.kappa[i];
  d.exp_wait = o.exp_wait;
  d.seed = (uint32_t)job_ID;
  d.salt = 0;
  d.delta = flow ? in.t_step_size : (double)(int)in.t_step_size;  // G(t) uses an integer interval (common/ensemble.cu:289)
  d.sample_mode = flow ? DSM_SAMPLE_FLOW
                       : in.equilibrium_type == 1 ? DSM_SAMPLE_GT
                       : in.equilibrium_type == 2 ? DSM_SAMPLE_MSD : DSM_SAMPLE_NONE;
  d.corr_p = o.corr_p;
  d.corr_m = o.corr_m;
  d.corr_levels = s;
  d.window = o.window;
  if (in.PD_flag) {
    d.partner_mode = DSM_PARTNER_GEX;
    for (int i = 0; i < 4; ++i) d.gex[i] = gex[i];
  }
  int ens = 0;
  if ((rc = dsm_add_ensemble(e, &d, &ens))) {
    std::cout << "error: " << dsm_last_error() << "\n";
    dsm_destroy(e);
    return rc;
  }
  HostState h;
  int cap = 0;
  dsm_state_cap(e, ens, &cap);
  h.resize(in.N_cha, cap);
  if (loadfile) {
    std::cout << "loading chain conformations from " << loadfile << "..";
    std::cout.flush();
    rc = dsm_load(e, ens, loadfile);
    if (rc == DSM_ERR_MISMATCH) std::cout << "ensemble size mismatch\n";
    else if (rc == DSM_ERR_IO) std::cout << "file error: " << g_io_err << "\n";
    else if (rc) std::cout << "error: " << dsm_last_error() << "\n";
    if (rc) { dsm_destroy(e); return rc; }
    std::cout << "done.\n";
  } else {
    std::cout << "generating chain conformations on device..";
    if ((rc = dsm_init_equilibrium(e, ens))) {
      std::cout << "error: " << dsm_last_error() << "\n";
      dsm_destroy(e);
      return rc;
    }
    std::cout << "done\n";
  }
  int32_t flag_i = 1;
  volatile int32_t prog = 0;
  auto t0 = std::chrono::steady_clock::now();
  double t_ref = 0.0;
  if (flow) {
    const std::string tau_name = filename_ID(dir, "tau", job_ID);
    std::ofstream tau_file(tau_name);
    std::cout << "output file: " << filename_ID("", "tau", job_ID) << '\n';
    std::cout << "performing time evolution for the ensemble..\n";
    std::cout << "time\tstress tensor(xx yy zz xy yz xz)\t<Lpp>\t<Z>\n";
    const int64_t n_out = (int64_t)std::ceil(in.simulation_time / in.t_step_size - 1e-12);
    int64_t printed = 0;
    const int64_t chunk = o.window > 0 ? o.window : 256;
    while (printed < n_out) {
      const int64_t target_rows = std::min(n_out, printed + chunk);
      const double t_target = (double)target_rows * in.t_step_size;
      if (run_flag && !*run_flag) { dsm_destroy(e); return -1; }
      flag_i = (run_flag && !*run_flag) ? 0 : 1;
      rc = dsm_advance(e, ens, t_target, o.mode, &flag_i, &prog, nullptr);
      if (rc) { std::cout << "error: " << dsm_last_error() << "\n"; dsm_destroy(e); return rc; }
      int64_t have = 0;
      dsm_flow_stress(e, ens, 0, 0, nullptr, &have);
      std::vector<double> rows((size_t)(have - printed) * 9);
      dsm_flow_stress(e, ens, printed, have - printed, rows.data(), &have);
      for (int64_t r = 0; r < (int64_t)rows.size() / 9; ++r) {
        const double* q = &rows[r * 9];
        const double cnt = q[8] > 0 ? q[8] : 1.0;
        float v[8];
        for (int k = 0; k < 8; ++k) v[k] = (float)(q[k] / cnt);
        std::ostringstream line;
        line << (double)(printed + r + 1) * in.t_step_size << '\t' << v[0] << ' ' << v[1] << ' ' << v[2] << ' '
             << v[3] << ' ' << v[4] << ' ' << v[5] << ' ' << v[6] << ' ' << v[7];
        if (!o.quiet) std::cout << line.str() << '\n';
        tau_file << line.str() << '\n';
        tau_file.flush();
      }
      printed = have;
      if (progress_bar) *progress_bar = (int)(100.0 * printed / std::max<int64_t>(1, n_out));
    }
    t_ref = (double)n_out * in.t_step_size;
    std::cout << "time evolution done.\n";
  } else if (in.equilibrium_type == 1 || in.equilibrium_type == 2) {
    // common/main_cuda.cu:142-155: G(t) (correlator type 0) or MSD(t) (type 1)
    const bool msd = in.equilibrium_type == 2;
    std::cout << (msd ? "MSD(t) calc...\n" : "G(t) calc...\n");
    std::cout.flush();
    const double T = in.simulation_time;
    const double step = std::max(1.0, d.delta) * (o.window > 0 ? o.window : 256) * 16;
    for (double tt = std::min(step, T);; tt = std::min(tt + step, T)) {
      if (run_flag && !*run_flag) { dsm_destroy(e); return -1; }
      rc = dsm_advance(e, ens, tt, o.mode, &flag_i, &prog, nullptr);
      if (rc) { std::cout << "error: " << dsm_last_error() << "\n"; dsm_destroy(e); return rc; }
      if (progress_bar) *progress_bar = (int)(100.0 * tt / T);
      if (!o.quiet) { std::cout << "\r" << (int)(100.0 * tt / T) << "%\t "; std::cout.flush(); }
      if (tt >= T) break;
    }
    t_ref = T;
    const int np_max = o.corr_p + (s - 1) * (o.corr_p - o.corr_p / o.corr_m);
    std::vector<double> lag(np_max), g(np_max);
    int np = 0;
    if ((rc = dsm_correlator(e, ens, np_max, lag.data(), g.data(), &np))) {
      std::cout << "error: " << dsm_last_error() << "\n";
      dsm_destroy(e);
      return rc;
    }
    // common/ensemble.cu:324-335: np - p + floor(T / (delta * m^(s-1))) points
    const int actual_np = (int)(np - o.corr_p +
                                std::floor(T / ((float)(int)in.t_step_size * std::pow((float)o.corr_m, (float)(s - 1)))));
    std::ofstream gf(filename_ID(dir, msd ? "MSD" : "G", job_ID));
    std::cout << "\n";
    for (int j = 0; j < actual_np && j < np; ++j) {
      const float tj = (float)(int64_t)lag[j];
      const float xj = (float)g[j];
      if (!o.quiet) std::cout << tj << '\t' << xj << '\n';
      gf << tj << '\t' << xj << '\n';
    }
  } else {
    std::cout << "There are no flow and no equilibrium quantity to calculate. Exiting... \n";
  }
  auto t1 = std::chrono::steady_clock::now();
  if (savefile || distr) {
    if ((rc = fetch(e, ens, h))) { dsm_destroy(e); return rc; }
  }
  if (savefile) {
    std::cout << "saving chain conformations to " << savefile << "..";
    if (dsm_save(e, ens, savefile, t_ref) != DSM_OK) std::cout << "file error\n";
    std::cout << "done.\n";
  }
  if (distr) {
    std::cout << "Saving distribution to file...";
    const std::string sfx = in.CD_flag ? "" : "_";
    save_Z_distribution(filename_ID(dir, "distr_Z" + sfx, 0), h);
    save_N_distribution(filename_ID(dir, "distr_N" + sfx, 0), h);
    save_Q_distribution(filename_ID(dir, "distr_Q" + sfx, 0), h);
  }
  std::cout << "Memory cleanup.. ";
  dsm_destroy(e);
  std::cout << "done.\n";
  std::cout << "Calculation time: " << std::chrono::duration<double, std::milli>(t1 - t0).count()
            << " milliseconds\n";
  return 0;
}

extern "C" int dsm_main(volatile const bool* run_flag, int job_ID, const char* savefile, const char* loadfile,
                        int device_ID, bool distr, volatile int* progress_bar) {
  return dsm_main_ex(run_flag, job_ID, savefile, loadfile, device_ID, distr, progress_bar, nullptr);
}
