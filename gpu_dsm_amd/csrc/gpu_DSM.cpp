// gpu_DSM.cpp -- command-line front-end with the argument surface of the
// reference CLI (ktaletsk/gpu_dsm CLI/main.cu:20-56, README.md:59-77):
//   gpu_DSM <job_ID> [-s file] [-l file] [-d device] [-distr]
// argv[1] is the job_ID (atoi; also the random seed).  Options this build adds
// (never stored in input.dat): --mode=plain|batched, --corr-p N, --corr-m N,
// --window N, --exp-wait, --quiet.
#include <cstdlib>
#include <cstring>
#include <iostream>

#include "dsm.h"

int main(int narg, char** arg) {
  bool run_flag = true;
  const char* savefile = nullptr;
  const char* loadfile = nullptr;
  bool distr = false;
  int device_ID = 0;
  int progress_bar = 0;
  int job_ID = 0;
  dsm_main_options opt{};
  opt.mode = DSM_MODE_BATCHED;
  opt.corr_p = 64;
  opt.corr_m = 8;
  int k = 1;
  while (k < narg) {
    if (k == 1) job_ID = std::atoi(arg[k]);
    if (std::strcmp(arg[k], "-s") == 0 && k + 1 < narg) {
      savefile = arg[k + 1];
      std::cout << "final chain conformations will be saved to " << savefile << '\n';
      ++k;
    } else if (std::strcmp(arg[k], "-l") == 0 && k + 1 < narg) {
      loadfile = arg[k + 1];
      ++k;
    } else if (std::strcmp(arg[k], "-d") == 0 && k + 1 < narg) {
      device_ID = std::atoi(arg[k + 1]);
      ++k;
    } else if (std::strcmp(arg[k], "-distr") == 0) {
      distr = true;
    } else if (std::strcmp(arg[k], "--mode=plain") == 0) {
      opt.mode = DSM_MODE_PLAIN;
    } else if (std::strcmp(arg[k], "--mode=batched") == 0) {
      opt.mode = DSM_MODE_BATCHED;
    } else if (std::strcmp(arg[k], "--corr-p") == 0 && k + 1 < narg) {
      opt.corr_p = std::atoi(arg[k + 1]);
      ++k;
    } else if (std::strcmp(arg[k], "--corr-m") == 0 && k + 1 < narg) {
      opt.corr_m = std::atoi(arg[k + 1]);
      ++k;
    } else if (std::strcmp(arg[k], "--window") == 0 && k + 1 < narg) {
      opt.window = std::atoi(arg[k + 1]);
      ++k;
    } else if (std::strcmp(arg[k], "--exp-wait") == 0) {
      opt.exp_wait = 1;
    } else if (std::strcmp(arg[k], "--quiet") == 0) {
      opt.quiet = 1;
    }
    ++k;
  }
  const int rc = dsm_main_ex(&run_flag, job_ID, savefile, loadfile, device_ID, distr, &progress_bar, &opt);
  return rc < 0 ? 1 : 0;
}
