// dsm_launch.h -- host-callable launchers for the kernels in dsm_kernels.hip.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "dsm_types.h"

namespace dsm {

// KMC kernels by chain geometry: 0 = row form (four chains per wave, Z <= 64),
// 1 = pair form (two chains per wave, Z <= 128), 2/3/4/5 = wave form (one chain per
// wave, Z <= 256/512/1024/2048): geometry g holds 64 << g positions per chain.  Persistent: grid =
// blocks of one wave; plain = one jump per chain.  lds_floats > 0 stages the rate
// tables (that many floats) in LDS.  a.ext selects the variant with MSD(t) tracking,
// polydisperse CD partners and the flow-mode shortcut.
constexpr int kNumGeoms = 6;
hipError_t launch_kmc(const KmcArgs& a, int geom, bool plain, int grid, hipStream_t st);
int kmc_max_blocks_per_cu(int geom, bool plain, int lds_floats, bool ext);
// Kernel geometries 0..5 hold 64 << g positions per chain (the work bins of a pass);
// geometry 7 is the two-group pair form, a second kernel for bin 2 (256 positions);
// shape 8, the row form with two register groups (128 positions, bin 1 of the merged
// row-form launch), names that shape for the probes only.
constexpr int kGeomPair2 = 7;
constexpr int kGeomRow2 = 8;
__host__ __device__ constexpr int kmc_geom_capacity(int geom) {
  return geom == kGeomPair2 ? 256 : geom == kGeomRow2 ? 128 : 64 << geom;
}
#define DSM_DECLARE_GEOM(G)                                                                   \
  hipError_t launch_kmc_geom_##G(const KmcArgs& a, bool plain, int grid, hipStream_t st); \
  int kmc_occupancy_geom_##G(bool plain, int lds_floats, bool ext);
DSM_DECLARE_GEOM(0)
DSM_DECLARE_GEOM(1)
DSM_DECLARE_GEOM(2)
DSM_DECLARE_GEOM(3)
DSM_DECLARE_GEOM(4)
DSM_DECLARE_GEOM(5)
DSM_DECLARE_GEOM(7)
#undef DSM_DECLARE_GEOM
// the row form (list `row`) and bin 1 (list `pair`: the row form with two register groups
// if row2, else the pair form) in one batched launch
hipError_t launch_kmc_rowpair(const KmcArgs& row, const KmcArgs& pair, bool row2, int grid, hipStream_t st);
int kmc_rowpair_occupancy(int lds_floats, bool ext, bool row2);
hipError_t launch_init_eq(const ChainArrays& ch, const EnsParams* ens, const float* tabs, const int* chains, int n,
                          hipStream_t st);
hipError_t launch_fill_u32(void* p, uint32_t v, int64_t n, hipStream_t st);
hipError_t launch_iota(int* p, int start, int n, hipStream_t st);
// the chains of src (entries < 0 skipped; src == null: chains first .. first+n-1) into
// dst in decreasing Z; *n_out = their count.  scratch17 (17 ints) must be zero.
hipError_t launch_z_order(const int* src, int first, int n, const int* Z, const int* stall, int* scratch17, int* dst,
                          int* n_out, hipStream_t st);
// cnt[0] = n0, cnt[1..n) = 0
hipError_t launch_init_counts(int* cnt, int n, int n0, hipStream_t st);
hipError_t launch_split_geoms(const int* src, int n, const int* Z, const int* stall, int* lists, int cap,
                              int* counts, int margin_div, hipStream_t st);
// out[0] += sum ev[chain0..chain0+n), out[1] += sum zsum[...]
hipError_t launch_sum_counters(const uint64_t* ev, const uint64_t* zsum, int chain0, int n, unsigned long long* out,
                               hipStream_t st);
hipError_t launch_set_ev_stop(const uint64_t* ev, uint64_t* ev_stop, int chain0, int n, uint64_t add,
                              hipStream_t st);
hipError_t launch_flow_reduce(const float* stage, int nch, int nrows, double* out,
                              hipStream_t st);
hipError_t launch_corr_update(const float* stage, double* D, double* A, double* C, const int* stall, int chain0,
                              int nch, int p, int m, int s, int nrows, const int* cnt, int type, hipStream_t st);
hipError_t launch_corr_reduce(const double* C, int ld, const int* stall, int chain0, int first, int count, int npts,
                              double* out_sum, int* out_healthy, hipStream_t st);
// ensemble stress of the current state: kStressBlocks partial sums of 9 doubles
// (xx yy zz xy yz zx Lpp Z count over the chains that have not stalled)
constexpr int kStressBlocks = 256;
struct KappaArg {
  float k[9];
};
hipError_t launch_chain_stress(const ChainArrays& ch, int chain0, int n, const float kappa[9], double* part,
                               hipStream_t st);
// window parameters (t_end, k0, k_end per ensemble) and correlator counters of a captured
// pass, from host-pinned memory into EnsParams and cnt_dst
hipError_t launch_stage_window(const WinParam* wp, int n_ens, EnsParams* P, const int* cnt_src, int* cnt_dst,
                               int n_cnt, hipStream_t st);
// invariant violations of chains chain0..chain0+n-1 added to out[0..6] (see dsm_check_state)
hipError_t launch_check_state(const ChainArrays& ch, int chain0, int n, double t_min, unsigned long long* out,
                              hipStream_t st);
hipError_t launch_copy_probe(const void* in, void* out, size_t n4, int grid, hipStream_t st);
hipError_t launch_scan_probe(const float* in, float* out, int n, int form, hipStream_t st);
hipError_t launch_com_probe(const float* in, float* out, int n, int geom, hipStream_t st);
hipError_t launch_math_probe(const float* in, float* out, int n, int which, hipStream_t st);
hipError_t launch_philox_probe(const uint32_t* ck, uint32_t* out, int n, hipStream_t st);

}  // namespace dsm
