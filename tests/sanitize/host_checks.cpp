// host_checks.cpp -- drives the host code of the library (the drop-in driver's input
// parsing, output writers and -s / -l files; the serving scheduler) over the host-only
// engine stand-in, for runs under AddressSanitizer / UndefinedBehaviorSanitizer.
// Usage: host_checks <scratch directory>; exit code 0 when every check passes.
#include <cstdio>
#include <cstdlib>
#include <fstream>
#include <string>
#include <vector>

#include "dsm.h"

static int failures = 0;
#define CHECK(c)                                                   \
  do {                                                             \
    if (!(c)) {                                                    \
      std::fprintf(stderr, "check failed at line %d: %s\n", __LINE__, #c); \
      ++failures;                                                  \
    }                                                              \
  } while (0)

static void write_input(const std::string& dir, int n, int nk, double kxy, int pd, int eq) {
  std::ofstream f(dir + "/input.dat");
  f << "1.0\n" << nk << "\n" << n << "\n0 " << kxy << " 0 0 0 0 0 0 0\n1\n" << pd << "\n" << eq << "\n1.0\n40.0\n";
}

int main(int argc, char** argv) {
  if (argc < 2) return 2;
  const std::string dir = argv[1];
  dsm_main_options o{};
  o.mode = DSM_MODE_BATCHED;
  o.corr_p = 16;
  o.corr_m = 2;
  o.quiet = 1;
  o.workdir = dir.c_str();
  bool run = true;
  int prog = 0;
  // flow run with -s and -distr, then G(t) and MSD runs, one with -l of that file
  write_input(dir, 50, 20, 0.01, 0, 0);
  const std::string save = dir + "/s.bin";
  CHECK(dsm_main_ex(&run, 0, save.c_str(), nullptr, 0, true, &prog, &o) == 0);
  CHECK(prog == 100);
  write_input(dir, 50, 20, 0.0, 0, 1);
  CHECK(dsm_main_ex(&run, 3, nullptr, save.c_str(), 0, false, &prog, &o) == 0);
  write_input(dir, 50, 20, 0.0, 0, 2);
  CHECK(dsm_main_ex(&run, 4, nullptr, nullptr, 0, true, &prog, &o) == 0);
  // mismatched size, truncated file, missing file, missing polydisp.dat, bad device
  write_input(dir, 51, 20, 0.0, 0, 1);
  CHECK(dsm_main_ex(&run, 0, nullptr, save.c_str(), 0, false, &prog, &o) == DSM_ERR_MISMATCH);
  {
    std::ifstream in(save, std::ios::binary);
    std::string bytes((std::istreambuf_iterator<char>(in)), std::istreambuf_iterator<char>());
    for (size_t cut : {(size_t)2, (size_t)30, bytes.size() / 2, bytes.size() - 1}) {
      std::ofstream out(dir + "/cut.bin", std::ios::binary);
      out.write(bytes.data(), (std::streamsize)cut);
      out.close();
      write_input(dir, 50, 20, 0.0, 0, 1);
      const std::string cutp = dir + "/cut.bin";
      CHECK(dsm_main_ex(&run, 0, nullptr, cutp.c_str(), 0, false, &prog, &o) == DSM_ERR_IO);
    }
  }
  const std::string missing = dir + "/missing.bin";
  CHECK(dsm_main_ex(&run, 0, nullptr, missing.c_str(), 0, false, &prog, &o) == DSM_ERR_IO);
  write_input(dir, 50, 20, 0.0, 1, 1);
  CHECK(dsm_main_ex(&run, 0, nullptr, nullptr, 0, false, &prog, &o) == DSM_ERR_IO);
  write_input(dir, 50, 20, 0.0, 0, 1);
  CHECK(dsm_main_ex(&run, 0, nullptr, nullptr, 7, false, &prog, &o) == DSM_ERR_DEVICE);
  bool stop = false;
  CHECK(dsm_main_ex(&stop, 0, nullptr, nullptr, 0, false, &prog, &o) == -1);
  // the serving scheduler: plan and a full (fast) stream
  dsm_serve_config c{};
  c.n_requests = 60; c.arrival_rate = 5000.0; c.seed = 9;
  c.nk_min = 16; c.nk_max = 512; c.nk_step = 16;
  c.n_beta = 3; c.beta[0] = 0.5f; c.beta[1] = 1.0f; c.beta[2] = 2.0f;
  c.nch_min = 64; c.nch_max = 2048; c.flow_fraction = 0.5; c.rate_min = 1e-5; c.rate_max = 1e-2;
  c.horizon = 4.0; c.window = 1.0; c.cd = 1; c.hbm_fraction = 0.5; c.progress_s = 0.0;
  std::vector<double> arr(c.n_requests), rate(c.n_requests), lat(c.n_requests),
      summary((size_t)c.n_requests * DSM_SERVE_SUMMARY);
  std::vector<int32_t> nk(c.n_requests), nch(c.n_requests);
  std::vector<float> beta(c.n_requests);
  std::vector<int64_t> bytes(c.n_requests);
  CHECK(dsm_serve_plan(&c, c.n_requests, arr.data(), nk.data(), beta.data(), nch.data(), rate.data(), bytes.data()) == 0);
  for (int i = 1; i < c.n_requests; ++i) CHECK(arr[i] >= arr[i - 1]);
  dsm_serve_report r{};
  CHECK(dsm_serve_ex(0, &c, &r, lat.data(), summary.data()) == 0);
  CHECK(r.completed == c.n_requests);
  CHECK(r.p50_s <= r.p99_s && r.p99_s <= r.max_s + 1e-12);
  for (int i = 0; i < c.n_requests; ++i) CHECK(summary[(size_t)i * DSM_SERVE_SUMMARY] == (rate[i] > 0 ? 2.0 : 1.0));
  c.nk_max = 5000;  // outside the supported range
  CHECK(dsm_serve(0, &c, &r, nullptr) == DSM_ERR_ARG);
  std::printf("host checks: %d failures\n", failures);
  return failures ? 1 : 0;
}
