// dsm_noalloc.hip -- dsm_stress_noalloc: the ensemble stress of chain state held in the
// caller's device memory, with no allocation of any kind.
//
// It runs the kernel of dsm_stress (chain_stress_kernel, dsm_kernels.hip), whose
// per-block partial sums live, for dsm_stress, in a buffer the engine allocates at
// dsm_create.  Here they live in `scratch` below, a file-scope device array of fixed
// size: kStressBlocks x 9 doubles (18 KiB), held by every GPU the library runs on until
// the process ends.  The grid is fixed at kStressBlocks blocks whatever the chain count,
// so every chain count fits the array; each call overwrites all of it, so nothing needs
// clearing between calls.  Calls that overlap on one GPU share the array and race.
#include <hip/hip_runtime.h>

#include <mutex>

#include "dsm.h"
#include "dsm_launch.h"

static __device__ __align__(256) double scratch[dsm::kStressBlocks * 9];

namespace {

constexpr int kMaxDevices = 64;
double* g_scratch[kMaxDevices] = {};  // device address of `scratch` on each GPU
std::mutex g_scratch_mu;

// The device address of this GPU's copy of `scratch` (the host-side name is not a device
// address): looked up once per device, with that device current.
hipError_t scratch_on(int device, double** out) {
  std::lock_guard<std::mutex> lock(g_scratch_mu);
  if (!g_scratch[device]) {
    void* p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return e;
    g_scratch[device] = (double*)p;
  }
  *out = g_scratch[device];
  return hipSuccess;
}

}  // namespace

extern "C" int dsm_stress_noalloc(int32_t device_id, const dsm_device_chains* c, int32_t n_chains,
                                  const float* kappa, double* out, int32_t* healthy) {
  if (!c || !out || n_chains < 0 || !c->Z || !c->stall || !c->base || !c->dt_prev || !c->Qx || !c->Qy || !c->Qz ||
      !c->N)
    return DSM_ERR_ARG;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || device_id < 0 || device_id >= ndev || device_id >= kMaxDevices)
    return DSM_ERR_DEVICE;
  if (hipSetDevice(device_id) != hipSuccess) return DSM_ERR_DEVICE;
  double* part = nullptr;
  if (scratch_on(device_id, &part) != hipSuccess) return DSM_ERR_DEVICE;
  dsm::ChainArrays ch{};
  ch.Z = const_cast<int*>(c->Z);
  ch.stall = const_cast<int*>(c->stall);
  ch.base = const_cast<int64_t*>(c->base);
  ch.dt_prev = const_cast<float*>(c->dt_prev);
  ch.Qx = const_cast<float*>(c->Qx);
  ch.Qy = const_cast<float*>(c->Qy);
  ch.Qz = const_cast<float*>(c->Qz);
  ch.N = const_cast<int*>(c->N);
  const float zero[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  if (dsm::launch_chain_stress(ch, 0, n_chains, kappa ? kappa : zero, part, nullptr) != hipSuccess)
    return DSM_ERR_DEVICE;
  double h[dsm::kStressBlocks * 9];
  if (hipMemcpy(h, part, sizeof(h), hipMemcpyDeviceToHost) != hipSuccess) return DSM_ERR_DEVICE;
  double sum[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
  for (int b = 0; b < dsm::kStressBlocks; ++b)
    for (int q = 0; q < 9; ++q) sum[q] += h[b * 9 + q];
  for (int q = 0; q < 8; ++q) out[q] = sum[8] > 0.0 ? sum[q] / sum[8] : 0.0;
  if (healthy) *healthy = (int32_t)sum[8];
  return DSM_OK;
}
