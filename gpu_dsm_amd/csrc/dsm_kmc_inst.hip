// dsm_kmc_inst.hip -- instantiates the KMC kernels for one chain geometry.
// Built once per geometry (-DDSM_GEOM=0..6) so the variants compile in parallel:
//   0: row form, four chains per wave, Z <= 64
//   1: pair form, two chains per wave, Z <= 128
//   2, 3, 4, 5: wave form, one chain per wave, Z <= 256, 512, 1024, 2048 (the last keeps
//   part of its state in scratch memory; it carries the rare chains with Z > 1024)
//   6: bins 0 and 1 of a concurrent pass in one launch: the row form with one register
//      group (Z < 64), then with two (64 <= Z < 128)
//   7: pair form with two register groups: two chains per wave, Z <= 256 (the bulk of
//      128 <= Z < 240 in concurrent passes)
#include "dsm_kmc.h"
#include "dsm_launch.h"

#ifndef DSM_GEOM
#error "DSM_GEOM must be defined"
#endif
#define DSM_CAT2(a, b) a##b
#define DSM_CAT(a, b) DSM_CAT2(a, b)

namespace dsm {

#if DSM_GEOM == 6
namespace {
constexpr int kMaxLdsFloats = 65536 / 4;
template <bool ROW2>
bool tables_fit(int lds_floats) { return lds_floats > 0 && row_pair_stage<ROW2>() + lds_floats <= kMaxLdsFloats; }

template <bool LT, bool EXT, bool ROW2>
hipError_t launch_rp(const KmcArgs& row, const KmcArgs& pair, int grid, hipStream_t st) {
  const size_t lds = sizeof(float) * ((size_t)row_pair_stage<ROW2>() + (LT ? (size_t)row.lds_floats : 0));
  hipLaunchKernelGGL((kmc_rowpair_kernel<LT, EXT, ROW2>), dim3(grid), dim3(64), lds, st, row, pair);
  return hipGetLastError();
}

template <bool LT, bool EXT, bool ROW2>
int occupancy_rp(int lds_floats) {
  int nb = 0;
  const size_t lds = sizeof(float) * ((size_t)row_pair_stage<ROW2>() + (LT ? (size_t)lds_floats : 0));
  const hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kmc_rowpair_kernel<LT, EXT, ROW2>, 64, lds);
  return e == hipSuccess ? nb : 0;
}

template <bool ROW2>
hipError_t launch_rp_any(const KmcArgs& row, const KmcArgs& pair, int grid, hipStream_t st) {
  const bool lt = tables_fit<ROW2>(row.lds_floats);
  if (row.ext)
    return lt ? launch_rp<true, true, ROW2>(row, pair, grid, st) : launch_rp<false, true, ROW2>(row, pair, grid, st);
  return lt ? launch_rp<true, false, ROW2>(row, pair, grid, st) : launch_rp<false, false, ROW2>(row, pair, grid, st);
}

template <bool ROW2>
int occupancy_rp_any(int lds_floats, bool ext) {
  const bool lt = tables_fit<ROW2>(lds_floats);
  if (ext) return lt ? occupancy_rp<true, true, ROW2>(lds_floats) : occupancy_rp<false, true, ROW2>(0);
  return lt ? occupancy_rp<true, false, ROW2>(lds_floats) : occupancy_rp<false, false, ROW2>(0);
}
}  // namespace

hipError_t launch_kmc_rowpair(const KmcArgs& row, const KmcArgs& pair, bool row2, int grid, hipStream_t st) {
  if (grid <= 0) return hipSuccess;
  return row2 ? launch_rp_any<true>(row, pair, grid, st) : launch_rp_any<false>(row, pair, grid, st);
}

int kmc_rowpair_occupancy(int lds_floats, bool ext, bool row2) {
  return row2 ? occupancy_rp_any<true>(lds_floats, ext) : occupancy_rp_any<false>(lds_floats, ext);
}

#else

namespace {
constexpr int kCPW = DSM_GEOM == 0 ? 4 : ((DSM_GEOM == 1 || DSM_GEOM == 7) ? 2 : 1);
constexpr int kG = DSM_GEOM <= 2 ? 1 : ((DSM_GEOM == 3 || DSM_GEOM == 7) ? 2 : (DSM_GEOM == 4 ? 4 : 8));
static_assert(Geo<kCPW, kG>::CAP == kmc_geom_capacity(DSM_GEOM), "positions per chain of the geometry");
// LDS a workgroup may hold: the rate tables are read from global memory when the staging
// arrays and the table copy would not fit together
constexpr int kMaxLdsFloats = 65536 / 4;
bool tables_fit(int lds_floats) { return lds_floats > 0 && stage_floats<kCPW, kG>() + lds_floats <= kMaxLdsFloats; }

template <bool LT, bool PLAIN, bool EXT>
hipError_t launch_one(const KmcArgs& a, int grid, hipStream_t st) {
  const size_t lds = sizeof(float) * ((size_t)stage_floats<kCPW, kG>() + (LT ? (size_t)a.lds_floats : 0));
  hipLaunchKernelGGL((kmc_kernel<kCPW, kG, LT, PLAIN, EXT>), dim3(grid), dim3(64), lds, st, a);
  return hipGetLastError();
}

template <bool LT, bool PLAIN, bool EXT>
int occupancy_one(int lds_floats) {
  int nb = 0;
  const size_t lds = sizeof(float) * ((size_t)stage_floats<kCPW, kG>() + (LT ? (size_t)lds_floats : 0));
  const hipError_t e =
      hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, kmc_kernel<kCPW, kG, LT, PLAIN, EXT>, 64, lds);
  return e == hipSuccess ? nb : 0;
}

template <bool EXT>
hipError_t launch_variant(const KmcArgs& a, bool plain, int grid, hipStream_t st) {
  const bool lt = tables_fit(a.lds_floats);
  if (plain) return lt ? launch_one<true, true, EXT>(a, grid, st) : launch_one<false, true, EXT>(a, grid, st);
  return lt ? launch_one<true, false, EXT>(a, grid, st) : launch_one<false, false, EXT>(a, grid, st);
}

template <bool EXT>
int occupancy_variant(bool plain, int lds_floats) {
  const bool lt = tables_fit(lds_floats);
  if (plain) return lt ? occupancy_one<true, true, EXT>(lds_floats) : occupancy_one<false, true, EXT>(0);
  return lt ? occupancy_one<true, false, EXT>(lds_floats) : occupancy_one<false, false, EXT>(0);
}
}  // namespace

hipError_t DSM_CAT(launch_kmc_geom_, DSM_GEOM)(const KmcArgs& a, bool plain, int grid, hipStream_t st) {
  if (grid <= 0) return hipSuccess;
  return a.ext ? launch_variant<true>(a, plain, grid, st) : launch_variant<false>(a, plain, grid, st);
}

int DSM_CAT(kmc_occupancy_geom_, DSM_GEOM)(bool plain, int lds_floats, bool ext) {
  return ext ? occupancy_variant<true>(plain, lds_floats) : occupancy_variant<false>(plain, lds_floats);
}
#endif

}  // namespace dsm
