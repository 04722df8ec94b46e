// dsm_device.h -- device-side building blocks for the DSM ensemble KMC kernels (gfx950).
//
// Everything here is FP32/integer arithmetic with a fixed operation order so the
// results are reproducible bit-for-bit by a plain C restatement on the host:
//   * Philox4x32-10 counter-based generator (Salmon et al., SC'11),
//   * exp: Cody-Waite reduction by ln2 + degree-7 Taylor polynomial (fmaf Horner),
//   * log: exponent split to m in [sqrt(1/2), sqrt(2)) + atanh series to s^11,
//   * sin/cos(2*pi*v): exact quadrant reduction on v + Taylor polynomials,
//   * a wave64 inclusive scan built from DPP row_shr:1,2,4,8 + row_bcast:15/31.
// IEEE division and square root are correctly rounded (hipcc default) and the
// translation unit is compiled with -ffp-contract=off.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

namespace dsm {

// Philox counter word 3 ("purpose").
constexpr uint32_t kPurposeStep = 0u;    // x0: pick, x1: exponential wait, x2: CD lifetime
constexpr uint32_t kPurposeGauss = 1u;   // Box-Muller normals for created strands
constexpr uint32_t kPurposeInitCut = 16u;
constexpr uint32_t kPurposeInitQ = 17u;
constexpr uint32_t kPurposeInitTau = 18u;
constexpr uint32_t kPurposeInitNk = 19u;
constexpr uint32_t kPurposeInitPartner = 20u;  // partner quantile of each initial link (top 16 bits)

enum EventKind : int {
  kEvNone = 0, kEvShiftP = 1, kEvShiftM = 2, kEvCDd = 3, kEvCDc = 4, kEvCDcTail = 5,
  kEvSDcHead = 6, kEvSDcTail = 7, kEvSDdHead = 8, kEvSDdTail = 9, kEvStall = 10
};

// 23 random bits: (k + 0.5) * 2^-23 is exact in FP32, so u lies in [2^-24, 1 - 2^-24]
__device__ __forceinline__ float u01(uint32_t x) { return ((float)(x >> 9) + 0.5f) * 0x1p-23f; }

__device__ __forceinline__ uint4 philox4x32_10(uint4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = make_uint4(hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0);
  }
  return c;
}

__device__ __forceinline__ float dexpf(float x) {
  if (x != x) return x;
  if (x > 88.0f) return __int_as_float(0x7f800000);
  if (x < -87.0f) return 0.0f;
  const float n = __builtin_rintf(x * 0x1.715476p+0f);
  float r = __builtin_fmaf(-n, 0x1.62e4p-1f, x);
  r = __builtin_fmaf(-n, 0x1.7f7d1cp-20f, r);
  float q = 0x1.a01a02p-13f;
  q = __builtin_fmaf(q, r, 0x1.6c16c2p-10f);
  q = __builtin_fmaf(q, r, 0x1.111112p-7f);
  q = __builtin_fmaf(q, r, 0x1.555556p-5f);
  q = __builtin_fmaf(q, r, 0x1.555556p-3f);
  q = __builtin_fmaf(q, r, 0x1p-1f);
  q = __builtin_fmaf(q, r, 1.0f);
  q = __builtin_fmaf(q, r, 1.0f);
  const int ni = (int)n;
  return q * __int_as_float((ni + 127) << 23);
}

__device__ __forceinline__ float dlogf(float x) {
  const uint32_t ix = __float_as_uint(x);
  int e = (int)(ix >> 23) - 127;
  uint32_t mb = (ix & 0x7fffffu) | 0x3f800000u;
  if (mb > 0x3fb504f3u) { mb -= 0x00800000u; e += 1; }
  const float m = __uint_as_float(mb);
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  float q = 0x1.745d18p-3f;
  q = __builtin_fmaf(z, q, 0x1.c71c72p-3f);
  q = __builtin_fmaf(z, q, 0x1.24924ap-2f);
  q = __builtin_fmaf(z, q, 0x1.99999ap-2f);
  q = __builtin_fmaf(z, q, 0x1.555556p-1f);
  const float lnm = __builtin_fmaf(s, z * q, s + s);
  const float fe = (float)e;
  return __builtin_fmaf(fe, 0x1.62e4p-1f, __builtin_fmaf(fe, 0x1.7f7d1cp-20f, lnm));
}

__device__ __forceinline__ void dsincos2pif(float v, float& sn, float& cs) {
  const float t = v - __builtin_rintf(v);
  const float qf = __builtin_rintf(4.0f * t);
  const float f = t - 0.25f * qf;
  const float z = f * f;
  float ps = 0x1.507834p+5f;
  ps = __builtin_fmaf(z, ps, -0x1.32d2ccp+6f);
  ps = __builtin_fmaf(z, ps, 0x1.466bc6p+6f);
  ps = __builtin_fmaf(z, ps, -0x1.4abbcep+5f);
  ps = __builtin_fmaf(z, ps, 0x1.921fb6p+2f);
  const float s = f * ps;
  float pc = -0x1.a6d1f2p+4f;
  pc = __builtin_fmaf(z, pc, 0x1.e1f506p+5f);
  pc = __builtin_fmaf(z, pc, -0x1.55d3c8p+6f);
  pc = __builtin_fmaf(z, pc, 0x1.03c1f0p+6f);
  pc = __builtin_fmaf(z, pc, -0x1.3bd3ccp+4f);
  const float c = __builtin_fmaf(z, pc, 1.0f);
  const int q = ((int)qf) & 3;
  sn = (q == 0) ? s : (q == 1) ? c : (q == 2) ? -s : -c;
  cs = (q == 0) ? c : (q == 1) ? -s : (q == 2) ? -c : s;
}

__device__ __forceinline__ float dpowf(float x, float y) { return dexpf(y * dlogf(x)); }

__device__ __forceinline__ void boxmuller(uint32_t a, uint32_t b, float& g0, float& g1) {
  const float r = sqrtf(-2.0f * dlogf(u01(a)));
  float s, c;
  dsincos2pif(u01(b), s, c);
  g0 = r * c;
  g1 = r * s;
}

// ---------------------------------------------------------------- wave64 helpers
template <int CTRL, int ROWMASK>
__device__ __forceinline__ float dpp_take(float v) {
  return __int_as_float(__builtin_amdgcn_update_dpp(0, __float_as_int(v), CTRL, ROWMASK, 0xF, false));
}

// Inclusive scan over the 64 lanes in the fixed tree order described above.
__device__ __forceinline__ float wave_scan(float v) {
  v = v + dpp_take<0x111, 0xF>(v);  // row_shr:1
  v = v + dpp_take<0x112, 0xF>(v);  // row_shr:2
  v = v + dpp_take<0x114, 0xF>(v);  // row_shr:4
  v = v + dpp_take<0x118, 0xF>(v);  // row_shr:8
  v = v + dpp_take<0x142, 0xA>(v);  // row_bcast:15 into rows 1,3
  v = v + dpp_take<0x143, 0xC>(v);  // row_bcast:31 into rows 2,3
  return v;
}

__device__ __forceinline__ float rdlane(float v, int lane) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}
__device__ __forceinline__ int rdlane(int v, int lane) { return __builtin_amdgcn_readlane(v, lane); }

__device__ __forceinline__ int lane_id() { return __lane_id(); }

// value of lane-1 (lane 0 receives `fill`)
__device__ __forceinline__ float from_prev_lane(float v, float fill) {
  const float t = __shfl_up(v, 1, 64);
  return lane_id() == 0 ? fill : t;
}
__device__ __forceinline__ int from_prev_lane(int v, int fill) {
  const int t = __shfl_up(v, 1, 64);
  return lane_id() == 0 ? fill : t;
}
// value of lane+1 (lane 63 receives `fill`)
__device__ __forceinline__ float from_next_lane(float v, float fill) {
  const float t = __shfl_down(v, 1, 64);
  return lane_id() == 63 ? fill : t;
}
__device__ __forceinline__ int from_next_lane(int v, int fill) {
  const int t = __shfl_down(v, 1, 64);
  return lane_id() == 63 ? fill : t;
}

}  // namespace dsm
