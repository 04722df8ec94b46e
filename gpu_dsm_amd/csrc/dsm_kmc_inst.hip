// dsm_kmc_inst.hip -- instantiates the KMC kernels for one register-slot count.
// Built once per S (-DDSM_SLOTS=1,2,4,8,16) so the large unrolled variants
// compile in parallel.
#include "dsm_kmc.h"
#include "dsm_launch.h"

#ifndef DSM_SLOTS
#error "DSM_SLOTS must be defined"
#endif
#define DSM_CAT2(a, b) a##b
#define DSM_CAT(a, b) DSM_CAT2(a, b)

namespace dsm {

hipError_t DSM_CAT(launch_kmc_slots_, DSM_SLOTS)(const KmcArgs& a, bool plain, int grid, hipStream_t st) {
  if (grid <= 0) return hipSuccess;
  if (plain) {
    const int blocks = (int)(((int64_t)grid * 64 + 255) / 256);
    hipLaunchKernelGGL(kmc_plain_kernel<DSM_SLOTS>, dim3(blocks), dim3(256), 0, st, a);
  } else {
    const size_t lds = (size_t)4 * a.lds_floats * sizeof(float);
    hipLaunchKernelGGL(kmc_batched_kernel<DSM_SLOTS>, dim3(grid), dim3(256), lds, st, a);
  }
  return hipGetLastError();
}

int DSM_CAT(kmc_occupancy_slots_, DSM_SLOTS)(bool plain, int lds_floats) {
  int nb = 0;
  const size_t lds = (size_t)4 * lds_floats * sizeof(float);
  const hipError_t e = plain ? hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kmc_plain_kernel<DSM_SLOTS>, 256, 0)
                             : hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kmc_batched_kernel<DSM_SLOTS>, 256, lds);
  return e == hipSuccess ? nb : 0;
}

}  // namespace dsm
