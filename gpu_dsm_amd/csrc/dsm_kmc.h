// dsm_kmc.h -- the KMC jump kernels, templated on the chain geometry.
// Included by dsm_kmc_inst.hip (one translation unit per geometry).
//
// Layout.  A chain's positions (strand i, and the pair/link between strands i and
// i+1) sit four to a lane: position p = 64*b + 4*l + k is slot k of lane l of the
// 16-lane row that holds block b of the chain.  Two shapes share that layout:
//   row form  (CPW = 4): one chain per 16-lane row, four chains per wave, Z <= 64;
//   wave form (CPW = 1): one chain per wave, block b in row b%4 of register group
//                        b/4, Z <= 256*G.
// Every FP32 total (rate sum W, stress samples) uses one canonical tree: serial over a
// lane's four slots, Hillis-Steele over the 16 lanes of a block, serial carry from
// block to block.  A chain therefore evolves identically in either shape.
//
// A jump (ktaletsk/gpu_dsm common/eq_ensemble_kernel.cu:242-663,
// common/ensemble_kernel.cu:188-502) is
//   1. cached pair rates at positions 0..Z-2, the five chain-end rates summed at Z-1,
//   2. the canonical scan -> W, dt = 1/W, FP64 clock += dt,
//   3. pr = W*u with u from Philox(seed, chain, jump); the pair is the first position
//      whose inclusive sum exceeds pr; the sub-event follows the reference order CDd,
//      SD+, SD-, CDc (pairs) / CDc tail, SDc head, SDc tail, SDd head, SDd tail,
//   4. the effect is applied to the register-resident chain (an N shift, or a
//      one-position gap opened/closed in every strand array plus patches), and
//   5. the cached rates of the (at most three) pairs next to the change are
//      recomputed in parallel, one pair per lane, from an LDS staging copy.
//
// Chain-uniform values (Z, clock, plan of the jump) are per-lane registers that hold
// the same value on every lane of the chain's segment, so the chain logic is plain
// per-lane code: rows that take different event branches are masked by exec, and a
// branch no row takes is skipped.  Cross-lane steps (DPP scans, swizzles,
// permutes) always run with the full wave active.
#pragma once
#include "dsm_device.h"
#include "dsm_types.h"

namespace dsm {

typedef __attribute__((address_space(3))) float lds_f;
typedef __attribute__((address_space(3))) int lds_i;
typedef float vf4 __attribute__((ext_vector_type(4)));
typedef int vi4 __attribute__((ext_vector_type(4)));
typedef __attribute__((address_space(3))) vf4 lds_f4;
typedef __attribute__((address_space(3))) vi4 lds_i4;

// ------------------------------------------------------------------ tables
// FP32 tables of one ensemble, indexed by the integer strand length N:
// SP=.75/(N(N+1)), SM=.75/(N(N-1)), PP=(N/(N+1))^.75, PM=(N/(N-1))^.75, EC, ED, ED2
// (end rates), T3=3/N, CC=c(N-1); then FR[x]=4/x, x=2(f_i+f_j).  Contiguous, so one
// base pointer and the row length n = nmax+2 address all of them.
template <bool LDS> struct TabPtr { typedef const __attribute__((address_space(1))) float* type; };
template <> struct TabPtr<true> { typedef const lds_f* type; };

// Table reads.  Global memory: a 32-bit byte offset from the wave-uniform base, so the
// read takes the scalar-base form (one shift per index instead of a sign extension and
// a 64-bit add); every index is a length or length sum, never negative.
__device__ __forceinline__ float tab_ld(const lds_f* b, int i) { return b[i]; }
__device__ __forceinline__ float tab_ld(const __attribute__((address_space(1))) float* b, int i) {
  typedef const __attribute__((address_space(1))) char* gbytes;
  return *(const __attribute__((address_space(1))) float*)((gbytes)b + (unsigned)i * 4u);
}

template <bool LDS>
struct Tabs {
  typename TabPtr<LDS>::type b;
  int n;
  __device__ __forceinline__ float at(int i) const { return tab_ld(b, i); }
  __device__ __forceinline__ float SP(int N) const { return tab_ld(b, N); }
  __device__ __forceinline__ float SM(int N) const { return tab_ld(b, n + N); }
  __device__ __forceinline__ float PP(int N) const { return tab_ld(b, 2 * n + N); }
  __device__ __forceinline__ float PM(int N) const { return tab_ld(b, 3 * n + N); }
  __device__ __forceinline__ float T3(int N) const { return tab_ld(b, 7 * n + N); }
  __device__ __forceinline__ float CC(int N) const { return tab_ld(b, 8 * n + N); }
  __device__ __forceinline__ float FR(int x) const { return tab_ld(b, 9 * n + x); }
};

// CD lifetime samplers returning the rate 1/tau (common/gpu_random.cu:143-148)
__device__ __forceinline__ float rate_ft(const EnsParams* P, float u) {
  return u < P->s_one_minus_g ? dpowf(__builtin_fmaf(u, P->s_At, P->s_Dt), P->s_Ct) : P->s_taud_inv;
}
__device__ __forceinline__ float rate_fdt(const EnsParams* P, float u) {
  return u < P->s_Bdt ? dpowf(__builtin_fmaf(u, P->s_Adt, P->s_Ddt), P->s_Cdt) : P->s_taud_inv;
}
// the same with the constants of partner quantile `idx` (polydisperse background,
// common/gpu_random.cu:212-228): c = s_At s_Ct s_Dt s_taud_inv s_one_minus_g (f_t) or
// s_Adt s_Bdt s_Cdt s_Ddt s_taud_inv (f_d_t)
__device__ __forceinline__ float rate_ft_partner(const float* pd, int idx, float u) {
  const float* c = pd + 5 * idx;
  return u < c[4] ? dpowf(__builtin_fmaf(u, c[0], c[2]), c[1]) : c[3];
}
__device__ __forceinline__ float rate_fdt_partner(const float* pd, int idx, float u) {
  const float* c = pd + 5 * (kPartnerK + idx);
  return u < c[1] ? dpowf(__builtin_fmaf(u, c[0], c[3]), c[2]) : c[4];
}

// Sliding-dynamics shift rates of the pair (i, j=i+1) and the pair's total with its
// CD terms (common/eq_ensemble_kernel.cu:97-126).
template <bool LDS>
__device__ __forceinline__ void pair_rate(const Tabs<LDS>& T, int cd, int Ni, int Nj, float q2i, float q2j,
                                          float we, bool endi, bool endj, float& wp, float& wm, float& R) {
  wp = 0.0f;
  wm = 0.0f;
  if (Nj > 1) {
    const float a = q2i * T.SP(Ni);
    const float b = q2j * T.SM(Nj);
    const int x = (endi ? 4 * Ni + 1 : 2 * Ni) + (endj ? 4 * Nj - 1 : 2 * Nj);
    const float pf = (T.FR(x) * (endi ? 1.0f : T.PP(Ni))) * (endj ? 1.0f : T.PM(Nj));
    wp = pf * dexpf(a - b);
  }
  if (Ni > 1) {
    const float a = q2i * T.SM(Ni);
    const float b = q2j * T.SP(Nj);
    const int x = (endi ? 4 * Ni - 1 : 2 * Ni) + (endj ? 4 * Nj + 1 : 2 * Nj);
    const float pf = (T.FR(x) * (endi ? 1.0f : T.PM(Ni))) * (endj ? 1.0f : T.PP(Nj));
    wm = pf * dexpf(b - a);
  }
  R = cd ? (wp + wm) + (we + T.CC(Ni)) : (wp + wm);
}

__device__ __forceinline__ float q2_of(float x, float y, float z) {
  return __builtin_fmaf(z, z, __builtin_fmaf(y, y, x * x));
}

__device__ __forceinline__ int split_newn(float r, int N, float wc) {
  int newn = (int)floorf(0.5f + (r * (float)(N - 2)) / wc) + 1;
  newn = newn < 1 ? 1 : newn;
  newn = newn > N - 1 ? N - 1 : newn;
  return newn;
}

// ------------------------------------------------------------------ geometry
template <int CPW, int G>
struct Geo {
  static_assert((CPW == 4 && (G == 1 || G == 2)) || (CPW == 2 && (G == 1 || G == 2)) ||
                    (CPW == 1 && (G == 1 || G == 2 || G == 4 || G == 8)),
                "unsupported chain geometry");
  static constexpr int R = 4 * G;                        // registers per strand array
  static constexpr int SEG = 64 / CPW;                   // lanes per chain
  static constexpr int GS = 4 * SEG;                     // positions per register group of a chain
  static constexpr int CAP = GS * G;                     // positions per chain
  static constexpr int STG = 256 * G;                    // floats per staging array (all chains of a wave)
};

// LDS staging arrays of a wave (each STG floats; chain positions at srow + p)
enum StageArray : int { kStN = 0, kStQ2, kStW, kStWP, kStWM, kStRR, kStArrays };
// The row form also stages the scan's inclusive sums (array kStG): its sub-event reads
// G[p-1] there, cheaper for four chains per wave than the pair and wave forms' slot
// select and cross-lane read, and its seven arrays still leave LDS for three waves per
// SIMD (the G = 2 and two-group pair kernels, at 512 floats per array, would get 11
// waves per CU instead of 12 with a seventh).
constexpr int kStG = kStArrays;
template <int CPW, int G>
constexpr bool kStageG = CPW == 4 && G == 1;

template <int CPW>
struct LaneC {
  int lane;  // 0..63
  int sl;    // lane within the chain's segment
  int sb;    // first lane of the segment
  int pos0;  // position of slot 0 of register group 0
  int srow;  // staging offset of the chain's position 0
  __device__ __forceinline__ LaneC() {
    lane = lane_id();
    sl = lane & (64 / CPW - 1);
    sb = lane & ~(64 / CPW - 1) & 63;
    pos0 = 4 * sl;
    srow = 4 * sb;
  }
  // register group g = j/4 holds positions 4*SEG*g .. (the chain's segment, four slots per lane)
  __device__ __forceinline__ int pos(int j) const { return (256 / CPW) * (j >> 2) + pos0 + (j & 3); }
};

// ------------------------------------------------------------------ cross-lane steps
template <int CTRL>
__device__ __forceinline__ int dppi(int v, int old) {
  return __builtin_amdgcn_update_dpp(old, v, CTRL, 0xF, 0xF, false);
}
template <int CTRL>
__device__ __forceinline__ float dppf(float v, float old) {
  return __int_as_float(__builtin_amdgcn_update_dpp(__float_as_int(old), __float_as_int(v), CTRL, 0xF, 0xF, false));
}
constexpr int kRowShr1 = 0x111, kRowShr2 = 0x112, kRowShr4 = 0x114, kRowShr8 = 0x118, kRowShl1 = 0x101;
constexpr int kWaveShl1 = 0x130, kWaveShr1 = 0x138;

// value of the previous / next lane of the chain's segment (0 outside the segment)
template <int CPW, typename T>
__device__ __forceinline__ T seg_prev(T v);
template <int CPW, typename T>
__device__ __forceinline__ T seg_next(T v);
template <> __device__ __forceinline__ float seg_prev<4, float>(float v) { return dppf<kRowShr1>(v, 0.0f); }
template <> __device__ __forceinline__ float seg_prev<1, float>(float v) { return dppf<kWaveShr1>(v, 0.0f); }
template <> __device__ __forceinline__ int seg_prev<4, int>(int v) { return dppi<kRowShr1>(v, 0); }
template <> __device__ __forceinline__ int seg_prev<1, int>(int v) { return dppi<kWaveShr1>(v, 0); }
template <> __device__ __forceinline__ float seg_next<4, float>(float v) { return dppf<kRowShl1>(v, 0.0f); }
template <> __device__ __forceinline__ float seg_next<1, float>(float v) { return dppf<kWaveShl1>(v, 0.0f); }
template <> __device__ __forceinline__ int seg_next<4, int>(int v) { return dppi<kRowShl1>(v, 0); }
template <> __device__ __forceinline__ int seg_next<1, int>(int v) { return dppi<kWaveShl1>(v, 0); }
// two chains per wave (32-lane segments): the wave shift crosses rows inside a segment;
// the lane at the segment's start (end) takes 0 instead of its neighbour's value
template <> __device__ __forceinline__ float seg_prev<2, float>(float v) {
  const float x = dppf<kWaveShr1>(v, 0.0f);
  return (lane_id() & 31) == 0 ? 0.0f : x;
}
template <> __device__ __forceinline__ int seg_prev<2, int>(int v) {
  const int x = dppi<kWaveShr1>(v, 0);
  return (lane_id() & 31) == 0 ? 0 : x;
}
template <> __device__ __forceinline__ float seg_next<2, float>(float v) {
  const float x = dppf<kWaveShl1>(v, 0.0f);
  return (lane_id() & 31) == 31 ? 0.0f : x;
}
template <> __device__ __forceinline__ int seg_next<2, int>(int v) {
  const int x = dppi<kWaveShl1>(v, 0);
  return (lane_id() & 31) == 31 ? 0 : x;
}

__device__ __forceinline__ float rdl(float v, int l) {
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
__device__ __forceinline__ int rdl(int v, int l) { return __builtin_amdgcn_readlane(v, l); }

// Wave form, across a 256-position register-group boundary.  wave_prev_x: lane l > 0
// gets cur from lane l-1, lane 0 gets prevg from lane 63; wave_next_x: lane l < 63 gets
// cur from lane l+1, lane 63 gets nextg from lane 0.  The per-lane choice of the source
// register is made before the permute, so no lane-dependent condition guards a
// cross-lane read (a read of lane 63's value under "lane == 0" is not to be simplified
// with facts about lane 0).
__device__ __forceinline__ int wave_prev_x(int cur, int prevg, int lane) {
  const int t = (lane == 63) ? prevg : cur;
  return __builtin_amdgcn_ds_bpermute(((lane + 63) & 63) << 2, t);
}
__device__ __forceinline__ float wave_prev_x(float cur, float prevg, int lane) {
  return __int_as_float(wave_prev_x(__float_as_int(cur), __float_as_int(prevg), lane));
}
__device__ __forceinline__ int wave_next_x(int cur, int nextg, int lane) {
  const int t = (lane == 0) ? nextg : cur;
  return __builtin_amdgcn_ds_bpermute(((lane + 1) & 63) << 2, t);
}
__device__ __forceinline__ float wave_next_x(float cur, float nextg, int lane) {
  return __int_as_float(wave_next_x(__float_as_int(cur), __float_as_int(nextg), lane));
}

// Across a register-group boundary inside a chain's segment (wave form, or the pair form
// with two groups): cross_prev gives segment lane sl > 0 `cur` of lane sl-1 and lane 0
// `prevg` of the segment's last lane; cross_next gives lane sl < SEG-1 `cur` of lane
// sl+1 and the segment's last lane `nextg` of its first lane.  The source register is
// chosen per lane before the permute, as in wave_prev_x / wave_next_x.
template <int CPW, typename T>
__device__ __forceinline__ T cross_prev(T cur, T prevg, const LaneC<CPW>& L) {
  if constexpr (CPW == 1) {
    return wave_prev_x(cur, prevg, L.lane);
  } else {
    constexpr int SEG = 64 / CPW;
    const T t = (L.sl == SEG - 1) ? prevg : cur;
    int b;
    __builtin_memcpy(&b, &t, 4);
    b = __builtin_amdgcn_ds_bpermute((L.sb + ((L.sl + SEG - 1) & (SEG - 1))) << 2, b);
    T r;
    __builtin_memcpy(&r, &b, 4);
    return r;
  }
}
template <int CPW, typename T>
__device__ __forceinline__ T cross_next(T cur, T nextg, const LaneC<CPW>& L) {
  if constexpr (CPW == 1) {
    return wave_next_x(cur, nextg, L.lane);
  } else {
    constexpr int SEG = 64 / CPW;
    const T t = (L.sl == 0) ? nextg : cur;
    int b;
    __builtin_memcpy(&b, &t, 4);
    b = __builtin_amdgcn_ds_bpermute((L.sb + ((L.sl + 1) & (SEG - 1))) << 2, b);
    T r;
    __builtin_memcpy(&r, &b, 4);
    return r;
  }
}

// value of v on segment lane `src` (per lane; 0 <= src < SEG)
template <int CPW>
__device__ __forceinline__ float seg_get(float v, int src, const LaneC<CPW>& L) {
  return __int_as_float(__builtin_amdgcn_ds_bpermute((L.sb + src) << 2, __float_as_int(v)));
}
template <int CPW>
__device__ __forceinline__ int seg_get(int v, int src, const LaneC<CPW>& L) {
  return __builtin_amdgcn_ds_bpermute((L.sb + src) << 2, v);
}

// lane 15 of each 16-lane row, broadcast over the row (bitmask swizzle and=0x10 or=0x0f)
__device__ __forceinline__ float row_last(float v) {
  return __int_as_float(__builtin_amdgcn_ds_swizzle(__float_as_int(v), 0x1F0));
}
__device__ __forceinline__ int row_last(int v) { return __builtin_amdgcn_ds_swizzle(v, 0x1F0); }
// lane IDX of each 32-lane half, broadcast over the half (bitmask swizzle and=0 or=IDX):
// the last lanes of a pair-form chain's two rows without v_readlane and a select
template <int IDX>
__device__ __forceinline__ int half_bcast(int v) {
  static_assert(IDX >= 0 && IDX < 32, "lane within a 32-lane half");
  return __builtin_amdgcn_ds_swizzle(v, IDX << 5);
}
template <int IDX>
__device__ __forceinline__ float half_bcast(float v) {
  return __int_as_float(half_bcast<IDX>(__float_as_int(v)));
}

// canonical block scan of one register group: inclusive values I[0..3]
__device__ __forceinline__ void block_scan4(float v0, float v1, float v2, float v3, float I[4]) {
  const float s0 = v0, s1 = s0 + v1, s2 = s1 + v2, s3 = s2 + v3;
  float H = s3;
  H = H + dppf<kRowShr1>(H, 0.0f);
  H = H + dppf<kRowShr2>(H, 0.0f);
  H = H + dppf<kRowShr4>(H, 0.0f);
  H = H + dppf<kRowShr8>(H, 0.0f);
  const float X = dppf<kRowShr1>(H, 0.0f);
  I[0] = X + s0;
  I[1] = X + s1;
  I[2] = X + s2;
  I[3] = X + s3;
}

// Canonical inclusive sums G of v over the chain (v is zero at positions >= Z) and
// the chain total (G at the last position of the last occupied block).
template <int CPW, int G>
__device__ __forceinline__ float chain_scan(const float (&v)[4 * G], float (&out)[4 * G], int Z,
                                            const LaneC<CPW>& L) {
  if constexpr (CPW == 4 && G == 2) {
    // blocks 0 and 1 of the chain are its row in register groups 0 and 1; block 1 adds
    // block 0's total
    float I1[4];
    block_scan4(v[0], v[1], v[2], v[3], out);
    block_scan4(v[4], v[5], v[6], v[7], I1);
    const float B0 = row_last(out[3]), B1 = row_last(I1[3]);
#pragma unroll
    for (int k = 0; k < 4; ++k) out[4 + k] = B0 + I1[k];
    return Z > 64 ? B0 + B1 : B0;
  } else if constexpr (CPW == 4) {
    block_scan4(v[0], v[1], v[2], v[3], out);
    return row_last(out[3]);
  } else if constexpr (CPW == 2 && G >= 2) {
    // chain s holds block 2g in row 2s and block 2g+1 in row 2s+1 of register group g;
    // the carry runs over the blocks in position order, as in the wave form
    const bool upper = (L.lane & 16) != 0;
    const int nb = ((Z - 1) >> 6) + 1;
    float carry = 0.0f, total = 0.0f;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      float I[4];
      block_scan4(v[4 * g], v[4 * g + 1], v[4 * g + 2], v[4 * g + 3], I);
      const float B0 = half_bcast<15>(I[3]), B1 = half_bcast<31>(I[3]);
      const float c0 = carry;
      carry = (g == 0) ? B0 : carry + B0;
      total = (2 * g == nb - 1) ? carry : total;
      const float c1 = carry;
      carry = carry + B1;
      total = (2 * g + 1 == nb - 1) ? carry : total;
#pragma unroll
      for (int k = 0; k < 4; ++k) out[4 * g + k] = upper ? c1 + I[k] : (g == 0 ? I[k] : c0 + I[k]);
    }
    return total;
  } else if constexpr (CPW == 2) {
    // blocks 0 and 1 of chain s are rows 2s and 2s+1; block 1 adds block 0's total
    float I[4];
    block_scan4(v[0], v[1], v[2], v[3], I);
    const float B0 = half_bcast<15>(I[3]), B1 = half_bcast<31>(I[3]);
    const bool upper = (L.lane & 16) != 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) out[k] = upper ? B0 + I[k] : I[k];
    return Z > 64 ? B0 + B1 : B0;
  } else if constexpr (G == 2) {
    // the serial carry over the four rows of each group: the row totals come from two
    // 32-lane swizzles (lower half: blocks 4g and 4g+1, upper half: 4g+2 and 4g+3), the
    // carry through block 4g+1 crosses to the upper half and the carry through block
    // 4g+3 to the next group by v_readlane; the additions are the same, in block order
    const int nb = ((__builtin_amdgcn_readfirstlane(Z) - 1) >> 6) + 1;
    const int row = L.lane >> 4;
    const bool up = L.lane >= 32, odd = (L.lane & 16) != 0;
    float carry = 0.0f, total = 0.0f;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      float I[4];
      block_scan4(v[4 * g], v[4 * g + 1], v[4 * g + 2], v[4 * g + 3], I);
      const float lo = half_bcast<15>(I[3]), hi = half_bcast<31>(I[3]);
      const float S0 = (g == 0) ? lo : carry + lo;  // through block 4g (lower half)
      const float S1 = S0 + hi;                     // through block 4g+1 (lower half)
      const float S1u = rdl(S1, 0);
      const float S2 = S1u + lo;                    // through block 4g+2 (upper half)
      const float S3 = S2 + hi;                     // through block 4g+3 (upper half)
      const float cin = up ? (odd ? S2 : S1u) : (odd ? S0 : carry);   // into the lane's row
      const float cout = up ? (odd ? S3 : S2) : (odd ? S1 : S0);      // through it
      const int last = nb - 1 - 4 * g;
      if (last >= 0 && last < 4) total = rdl(cout, 16 * last);
      carry = rdl(S3, 32);
#pragma unroll
      for (int k = 0; k < 4; ++k) out[4 * g + k] = (g == 0 && row == 0) ? I[k] : cin + I[k];
    }
    return total;
  } else {
    const int nb = ((__builtin_amdgcn_readfirstlane(Z) - 1) >> 6) + 1;
    const int row = L.lane >> 4;
    float carry = 0.0f, total = 0.0f;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      float I[4];
      block_scan4(v[4 * g], v[4 * g + 1], v[4 * g + 2], v[4 * g + 3], I);
      float cr = 0.0f;
#pragma unroll
      for (int r = 0; r < 4; ++r) {
        const int b = 4 * g + r;
        const float Bt = rdl(I[3], 16 * r + 15);
        cr = (row == r) ? carry : cr;
        carry = (b == 0) ? Bt : carry + Bt;
        if (b == nb - 1) total = carry;
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) out[4 * g + k] = (g == 0 && row == 0) ? I[k] : cr + I[k];
    }
    return total;
  }
}

template <int CPW, int G>
__device__ __forceinline__ float chain_total(const float (&v)[4 * G], int Z, const LaneC<CPW>& L) {
  float tmp[4 * G];
  return chain_scan<CPW, G>(v, tmp, Z, L);
}

// First position p < Z with G_p > pr, replicated over the segment; CAP when none.
template <int CPW, int G>
__device__ __forceinline__ int pick_position(const float (&Gv)[4 * G], float pr, int Z, const LaneC<CPW>& L) {
  constexpr int CAP = Geo<CPW, G>::CAP;
  if constexpr (CPW == 4 && G == 2) {
    // each lane's first slot over both groups (group 1 holds the later positions, so the
    // slots are visited from the last of group 1 down to the first of group 0), then one
    // row minimum: the same position as a minimum per group
    constexpr int kIdent = 0x7fffffff;
    int cand = CAP;
#pragma unroll
    for (int g = 1; g >= 0; --g) {
      const int b0 = 64 * g + L.pos0;
#pragma unroll
      for (int k = 3; k >= 0; --k) cand = (Gv[4 * g + k] > pr && b0 + k < Z) ? b0 + k : cand;
    }
    cand = min(cand, dppi<kRowShr1>(cand, kIdent));
    cand = min(cand, dppi<kRowShr2>(cand, kIdent));
    cand = min(cand, dppi<kRowShr4>(cand, kIdent));
    cand = min(cand, dppi<kRowShr8>(cand, kIdent));
    return row_last(cand);
  } else if constexpr (CPW == 4) {
    int cand = CAP;
#pragma unroll
    for (int k = 3; k >= 0; --k) cand = (Gv[k] > pr && L.pos0 + k < Z) ? L.pos0 + k : cand;
    // row minimum (sources before the row's lane 0 contribute the identity)
    constexpr int kIdent = 0x7fffffff;
    cand = min(cand, dppi<kRowShr1>(cand, kIdent));
    cand = min(cand, dppi<kRowShr2>(cand, kIdent));
    cand = min(cand, dppi<kRowShr4>(cand, kIdent));
    cand = min(cand, dppi<kRowShr8>(cand, kIdent));
    return row_last(cand);
  } else if constexpr (CPW == 2 && G >= 2) {
    // each lane's first slot over all groups (later groups hold later positions, so the
    // slots are visited from the last group down to the first), then one minimum over
    // the chain's two rows: the same position as a minimum per group
    constexpr int kIdent = 0x7fffffff;
    int cand = CAP;
#pragma unroll
    for (int g = G - 1; g >= 0; --g) {
      const int b0 = Geo<CPW, G>::GS * g + L.pos0;
#pragma unroll
      for (int k = 3; k >= 0; --k) cand = (Gv[4 * g + k] > pr && b0 + k < Z) ? b0 + k : cand;
    }
    cand = min(cand, dppi<kRowShr1>(cand, kIdent));
    cand = min(cand, dppi<kRowShr2>(cand, kIdent));
    cand = min(cand, dppi<kRowShr4>(cand, kIdent));
    cand = min(cand, dppi<kRowShr8>(cand, kIdent));
    return min(half_bcast<15>(cand), half_bcast<31>(cand));
  } else if constexpr (CPW == 2) {
    int cand = CAP;
#pragma unroll
    for (int k = 3; k >= 0; --k) cand = (Gv[k] > pr && L.pos0 + k < Z) ? L.pos0 + k : cand;
    constexpr int kIdent = 0x7fffffff;
    cand = min(cand, dppi<kRowShr1>(cand, kIdent));
    cand = min(cand, dppi<kRowShr2>(cand, kIdent));
    cand = min(cand, dppi<kRowShr4>(cand, kIdent));
    cand = min(cand, dppi<kRowShr8>(cand, kIdent));
    return min(half_bcast<15>(cand), half_bcast<31>(cand));
  } else {
    int res = CAP;
#pragma unroll
    for (int g = 0; g < G; ++g) {
      int kf = 4;
#pragma unroll
      for (int k = 3; k >= 0; --k) kf = (Gv[4 * g + k] > pr && 256 * g + L.pos0 + k < Z) ? k : kf;
      const unsigned long long m = __ballot(kf < 4);
      if (res == CAP && m) {
        const int ln = __ffsll((long long)m) - 1;
        res = 256 * g + 4 * ln + rdl(kf, ln);
      }
    }
    return res;
  }
}

// value of register array a at position p (p replicated over the segment, 0 <= p < CAP).
// Every register is permuted and the slot is selected afterwards, so the array is
// never indexed at run time (which would force it out of registers).
template <int CPW, int G, typename T>
__device__ __forceinline__ T value_at(const T (&a)[4 * G], int p, const LaneC<CPW>& L) {
  const int j = 4 * (p / Geo<CPW, G>::GS) + (p & 3);
  const int src = (p >> 2) & (Geo<CPW, G>::SEG - 1);
  T v = seg_get<CPW>(a[0], src, L);
#pragma unroll
  for (int q = 1; q < 4 * G; ++q) {
    const T x = seg_get<CPW>(a[q], src, L);
    v = (j == q) ? x : v;
  }
  return v;
}

// new[p] = (at < p <= Z) ? old[p-1] : old[p]   (at, Z replicated; at >= CAP: no-op)
template <int CPW, int G, typename T>
__device__ __forceinline__ void gap_open(T (&a)[4 * G], int at, int Z, const LaneC<CPW>& L) {
  T prev3[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    prev3[g] = (g > 0) ? cross_prev<CPW, T>(a[4 * g + 3], a[g > 0 ? 4 * g - 1 : 0], L)
                                   : seg_prev<CPW, T>(a[4 * g + 3]);
  }
#pragma unroll
  for (int j = 4 * G - 1; j >= 0; --j) {
    const int p = L.pos(j);
    const T src = (j & 3) ? a[j > 0 ? j - 1 : 0] : prev3[j >> 2];
    a[j] = (p > at && p <= Z) ? src : a[j];
  }
}
// new[p] = (at <= p < Z-1) ? old[p+1] : old[p]
template <int CPW, int G, typename T>
__device__ __forceinline__ void gap_close(T (&a)[4 * G], int at, int Z, const LaneC<CPW>& L) {
  T next0[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    next0[g] = (g + 1 < G) ? cross_next<CPW, T>(a[4 * g], a[4 * g + 4 < 4 * G ? 4 * g + 4 : 0], L)
                                       : seg_next<CPW, T>(a[4 * g]);
  }
#pragma unroll
  for (int j = 0; j < 4 * G; ++j) {
    const int p = L.pos(j);
    const T src = ((j & 3) != 3) ? a[j + 1 < 4 * G ? j + 1 : j] : next0[j >> 2];
    a[j] = (p >= at && p < Z - 1) ? src : a[j];
  }
}

// out[p] = a[p-1], out[0] = 0 (whole wave)
template <int CPW, int G>
__device__ __forceinline__ void shift_prev(const float (&a)[4 * G], float (&out)[4 * G], const LaneC<CPW>& L) {
#pragma unroll
  for (int g = 0; g < G; ++g) {
    out[4 * g] = (g > 0) ? cross_prev<CPW, float>(a[4 * g + 3], a[g > 0 ? 4 * g - 1 : 0], L)
                                     : seg_prev<CPW, float>(a[4 * g + 3]);
#pragma unroll
    for (int k = 1; k < 4; ++k) out[4 * g + k] = a[4 * g + k - 1];
  }
}

// ------------------------------------------------------------------ per-wave LDS staging
template <int CPW, int G>
struct Stage {
  lds_f* s;  // wave base
  int srow;
  __device__ __forceinline__ lds_f* arr(int a) const { return s + a * Geo<CPW, G>::STG + srow; }
  __device__ __forceinline__ float getf(int a, int p) const { return arr(a)[p]; }
  __device__ __forceinline__ int geti(int a, int p) const { return ((lds_i*)arr(a))[p]; }
  __device__ __forceinline__ void putf(int a, int p, float v) const { arr(a)[p] = v; }
  __device__ __forceinline__ void puti(int a, int p, int v) const { ((lds_i*)arr(a))[p] = v; }
  // this lane's four slots of register group g
  __device__ __forceinline__ void put4(int a, int g, int pos0, float v0, float v1, float v2, float v3) const {
    const vf4 v = {v0, v1, v2, v3};
    *(lds_f4*)(arr(a) + Geo<CPW, G>::GS * g + pos0) = v;
  }
  __device__ __forceinline__ void put4i(int a, int g, int pos0, int v0, int v1, int v2, int v3) const {
    const vi4 v = {v0, v1, v2, v3};
    *(lds_i4*)((lds_i*)arr(a) + Geo<CPW, G>::GS * g + pos0) = v;
  }
  __device__ __forceinline__ vf4 get4(int a, int g, int pos0) const {
    return *(lds_f4*)(arr(a) + Geo<CPW, G>::GS * g + pos0);
  }
  __device__ __forceinline__ vi4 get4i(int a, int g, int pos0) const {
    return *(lds_i4*)((lds_i*)arr(a) + Geo<CPW, G>::GS * g + pos0);
  }
};

// this lane's strand lengths N (all register groups) from the staging array
template <int CPW, int G>
__device__ __forceinline__ void get_n(const Stage<CPW, G>& S, const LaneC<CPW>& L, int (&n)[4 * G]) {
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const vi4 v = S.get4i(kStN, g, L.pos0);
    n[4 * g] = v.x; n[4 * g + 1] = v.y; n[4 * g + 2] = v.z; n[4 * g + 3] = v.w;
  }
}

// orders this wave's LDS writes before its later reads (cross-lane exchange)
__device__ __forceinline__ void wave_lds_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// ------------------------------------------------------------------ chain state
template <int G>
struct Strands {
  // (|Q|^2 of a strand is recomputed from qx, qy, qz where it is needed: the same FP32
  // expression every time, so no register copy of it is kept.  The strand lengths N live
  // in the LDS staging array kStN only: a shift changes two of them, which one lane
  // writes there, and every reader -- end rates, sub-event, pair rates, samples, the
  // write-back -- reads them from LDS.)
  float qx[4 * G], qy[4 * G], qz[4 * G], w[4 * G];
};

// Chain-uniform values, replicated over the chain's lanes.
struct Hdr {
  int c;          // chain index, -1: empty segment
  int inert;      // empty, stalled or over capacity at load, or overflowed since: no samples, no jumps
  int Z, nk, stall, local;
  uint32_t cid;
  int64_t next_sample;
  double t, t_next;
  uint64_t ev, zsum;
  int left;       // jumps left before the chain's ev_stop (saturated)
  float dt_prev, e1;
  int ovf, done, dlo, dhi;
};

// Per-wave ensemble values (every chain of a wave belongs to one ensemble).
struct EnsLocal {
  const EnsParams* P;
  int cd, flow, exp_wait, sample_mode, nchains, nmax, partner;
  uint32_t key0, key1;
  double delta, t_end;
  int64_t k0, k_end, stage_off;
  float beta;
};

__device__ __forceinline__ void load_ens(EnsLocal& E, const EnsParams* __restrict__ ps, int e) {
  const EnsParams& P = ps[e];
  E.P = &P;
  E.cd = P.cd; E.flow = P.flow; E.exp_wait = P.exp_wait; E.sample_mode = P.sample_mode;
  E.nchains = P.nchains; E.nmax = P.nmax; E.key0 = P.key0; E.key1 = P.key1; E.delta = P.delta;
  E.partner = P.partner;
  E.t_end = P.t_end; E.k0 = P.k0; E.k_end = P.k_end; E.stage_off = P.stage_off; E.beta = P.beta;
}

__device__ __forceinline__ double next_time(const EnsLocal& E, int64_t next_sample) {
  return next_sample < E.k_end ? (double)next_sample * E.delta : __longlong_as_double(0x7ff0000000000000ll);
}

// Philox outputs for SEG consecutive jumps of the chain, jump base+sl on segment lane sl.
struct RandCache {
  float sel, wait, tau, g0, g1, g2;
};

__device__ __forceinline__ void refill_rand(RandCache& R, uint64_t ev, int sl, uint32_t cid, const EnsLocal& E) {
  const uint64_t e = ev + (uint64_t)sl;
  const uint32_t lo = (uint32_t)e, hi = (uint32_t)(e >> 32);
  const uint4 x = philox4x32_10(make_uint4(lo, hi, cid, kPurposeStep), E.key0, E.key1);
  R.sel = u01(x.x); R.wait = u01(x.y); R.tau = u01(x.z);
  const uint4 g = philox4x32_10(make_uint4(lo, hi, cid, kPurposeGauss), E.key0, E.key1);
  float g3;
  boxmuller(g.x, g.y, R.g0, R.g1);
  boxmuller(g.z, g.w, R.g2, g3);
}

// ------------------------------------------------------------------ building blocks
// Stage |Q|^2 and the CD rates; q2 receives the |Q|^2 values for a refresh_all that
// follows.
template <int CPW, int G>
__device__ __forceinline__ void stage_qw(const Stage<CPW, G>& S, const Strands<G>& C, const LaneC<CPW>& L,
                                         float (&q2)[4 * G]) {
#pragma unroll
  for (int j = 0; j < 4 * G; ++j) q2[j] = q2_of(C.qx[j], C.qy[j], C.qz[j]);
#pragma unroll
  for (int g = 0; g < G; ++g) {
    S.put4(kStQ2, g, L.pos0, q2[4 * g], q2[4 * g + 1], q2[4 * g + 2], q2[4 * g + 3]);
    S.put4(kStW, g, L.pos0, C.w[4 * g], C.w[4 * g + 1], C.w[4 * g + 2], C.w[4 * g + 3]);
  }
}
template <int CPW, int G>
__device__ __forceinline__ void stage_qw(const Stage<CPW, G>& S, const Strands<G>& C, const LaneC<CPW>& L) {
  float q2[4 * G];
  stage_qw<CPW, G>(S, C, L, q2);
}

// The five chain-end rates (common/eq_ensemble_kernel.cu:340-392) from the staged
// strand lengths at both ends: CD creation in the tail strand, SD creation and
// destruction at the head and the tail.  They change only when the lengths at a chain
// end (N0, N1, N[Z-2], N[Z-1]) or Z do.  The pair and wave forms keep their sum in the
// rate cache at position Z-1 and recompute it after a jump that changed one of those
// (one to two chains per wave: a few percent of the jumps of a long chain).  The row
// form, whose four short chains touch an end in most steps, computes it at every jump
// and keeps the cache at Z-1 zero (measured on the 131,072-chain Nk = 100 melt: the row
// kernel 8.9 ms that way against 9.1 ms cached; the 262,144-chain polydisperse melt's
// window 98.5-99 ms against 102-104 ms with every form computing them per jump).
template <int CPW>
constexpr bool kEndCache = CPW <= 2;
struct EndRates {
  float cdcz, sdc1, sdcz, sdd1, sddz;
  __device__ __forceinline__ float total() const { return (((cdcz + sdc1) + sdcz) + sdd1) + sddz; }
};
template <int CPW, int G, bool LT>
__device__ __forceinline__ EndRates end_rates(const Stage<CPW, G>& S, const Tabs<LT>& T, int Z, float e1) {
  const int N0 = S.geti(kStN, 0), N1 = S.geti(kStN, 1);
  const int Nz = S.geti(kStN, Z - 1), Nz2 = S.geti(kStN, Z >= 2 ? Z - 2 : 0);
  const bool z2 = (Z == 2);
  const int ho = (N0 == 1) ? (z2 ? 6 : 5) * T.n + N1 : 4 * T.n + N0;
  const int to = (Nz == 1) ? (z2 ? 6 : 5) * T.n + Nz2 : 4 * T.n + Nz;
  const float h = T.at(ho), tt = T.at(to);
  EndRates r;
  r.cdcz = T.CC(Nz);
  if (Z == 1) {
    r.sdc1 = e1; r.sdcz = e1; r.sdd1 = 0.0f; r.sddz = 0.0f;
  } else {
    r.sdd1 = (N0 == 1) ? h : 0.0f;
    r.sdc1 = (N0 == 1) ? 0.0f : h;
    r.sddz = (Nz == 1) ? tt : 0.0f;
    r.sdcz = (Nz == 1) ? 0.0f : tt;
  }
  return r;
}

// Recompute every cached pair rate of the chain (after a load or a deformation):
// each lane evaluates the pairs of its own slots; the chain's first lane then puts the
// chain-end rates at Z-1 (after the other lanes' lengths are visible).  Segments without
// a live chain (empty, finished, stalled, overflowed) read nothing and cache zeros: their
// lengths may be a finished chain's, of another ensemble than the tables.
template <int CPW, int G, bool LT>
__device__ __forceinline__ void refresh_all(const Stage<CPW, G>& S, const Strands<G>& C, const float (&q2)[4 * G],
                                            const Tabs<LT>& T, int cd, int Z, float e1, bool live,
                                            const LaneC<CPW>& L) {
  int n[4 * G];
  get_n<CPW, G>(S, L, n);
  int nn[G];
  float qn[G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int gn = g + 1 < G ? g + 1 : g;
    const bool cross = g + 1 < G;
    nn[g] = cross ? cross_next<CPW, int>(n[4 * g], n[4 * gn], L) : seg_next<CPW, int>(n[4 * g]);
    qn[g] = cross ? cross_next<CPW, float>(q2[4 * g], q2[4 * gn], L) : seg_next<CPW, float>(q2[4 * g]);
  }
#pragma unroll
  for (int g = 0; g < G; ++g) {
    float wp[4], wm[4], rr[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int j = 4 * g + k;
      const int p = L.pos(j);
      const int Nj = (k < 3) ? n[j + (k < 3 ? 1 : 0)] : nn[g];
      const float q2j = (k < 3) ? q2[j + (k < 3 ? 1 : 0)] : qn[g];
      wp[k] = 0.0f; wm[k] = 0.0f; rr[k] = 0.0f;
      if (live && p <= Z - 2)
        pair_rate<LT>(T, cd, n[j], Nj, q2[j], q2j, C.w[j], p == 0, p + 1 == Z - 1, wp[k], wm[k], rr[k]);
    }
    S.put4(kStWP, g, L.pos0, wp[0], wp[1], wp[2], wp[3]);
    S.put4(kStWM, g, L.pos0, wm[0], wm[1], wm[2], wm[3]);
    S.put4(kStRR, g, L.pos0, rr[0], rr[1], rr[2], rr[3]);
  }
  if constexpr (kEndCache<CPW>) {
    wave_lds_sync();
    if (live) {
      const float eend = end_rates<CPW, G, LT>(S, T, Z, e1).total();
      if (L.sl == 0) S.putf(kStRR, Z - 1, eend);
    }
  }
}

// Recompute the cached rates of pairs lo..hi (at most three, clamped to 0..Z-2).
// Each pair is split over two lanes of the segment's first row running one
// instruction stream: lane j (0..2) evaluates W+ of pair lo+j and lane 8+j its W-
// (the same FP operations as pair_rate, with the tables and signs chosen per lane);
// lane j then adds the halves (W- arrives by DPP row_ror:8) and writes W+, W-, R.
// Lane 3 zeroes the rate cache at zero_at (>= 0) after a destruction: the old end
// position (pair and wave forms, whose cache holds the chain-end rates there) or the new
// one (row form).
template <int CPW, int G, bool LT>
__device__ __forceinline__ void refresh_pairs(const Stage<CPW, G>& S, const Tabs<LT>& T, int cd, int Z, int lo, int hi,
                                              int zero_at, const LaneC<CPW>& L) {
  const int h = (L.sl >> 3) & 1;  // 0: W+, 1: W-
  const int jl = L.sl & 7;
  const int q = lo + jl;
  const bool on = L.sl < 16 && jl < 3 && q <= hi && q >= 0 && q <= Z - 2;
  float part = 0.0f, we = 0.0f;
  int Ni = 0;
  if (on) {
    Ni = S.geti(kStN, q);
    const int Nj = S.geti(kStN, q + 1);
    const float q2i = S.getf(kStQ2, q), q2j = S.getf(kStQ2, q + 1);
    we = S.getf(kStW, q);
    const bool endi = q == 0, endj = q + 1 == Z - 1;
    const int n = T.n;
    if (h ? Ni > 1 : Nj > 1) {
      // W+: a = q2i SP(Ni), b = q2j SM(Nj), FR(x) PP(Ni) PM(Nj), exp(a - b)
      // W-: a = q2i SM(Ni), b = q2j SP(Nj), FR(x) PM(Ni) PP(Nj), exp(b - a)
      const float a = q2i * T.at((h ? n : 0) + Ni);
      const float b = q2j * T.at((h ? 0 : n) + Nj);
      const int x = (endi ? 4 * Ni + (h ? -1 : 1) : 2 * Ni) + (endj ? 4 * Nj + (h ? 1 : -1) : 2 * Nj);
      const float pf = (T.FR(x) * (endi ? 1.0f : T.at((h ? 3 : 2) * n + Ni))) *
                       (endj ? 1.0f : T.at((h ? 2 : 3) * n + Nj));
      part = pf * dexpf(h ? b - a : a - b);
    }
  }
  const float wm = dppf<0x128>(part, 0.0f);  // row_ror:8: lane j receives lane 8+j
  if (on && h == 0) {
    const float wp = part;
    const float R = cd ? (wp + wm) + (we + T.CC(Ni)) : (wp + wm);
    S.putf(kStWP, q, wp);
    S.putf(kStWM, q, wm);
    S.putf(kStRR, q, R);
  }
  if (L.sl == 3 && zero_at >= 0) S.putf(kStRR, zero_at, 0.0f);
}

// Gap op on the LDS-resident rate caches (wp, wm, rr).
template <int CPW, int G>
__device__ __forceinline__ void gap_lds(const Stage<CPW, G>& S, int dir, int at, int Z, const LaneC<CPW>& L) {
#pragma unroll
  for (int a = kStWP; a <= kStRR; ++a) {
    float v[4 * G];
#pragma unroll
    for (int g = 0; g < G; ++g) {
      const vf4 x = S.get4(a, g, L.pos0);
      v[4 * g] = x.x; v[4 * g + 1] = x.y; v[4 * g + 2] = x.z; v[4 * g + 3] = x.w;
    }
    if (dir > 0) gap_open<CPW, G>(v, at, Z, L);
    else gap_close<CPW, G>(v, at, Z, L);
#pragma unroll
    for (int g = 0; g < G; ++g) S.put4(a, g, L.pos0, v[4 * g], v[4 * g + 1], v[4 * g + 2], v[4 * g + 3]);
  }
}

// Gap op on the LDS-resident strand lengths.
template <int CPW, int G>
__device__ __forceinline__ void gap_n(const Stage<CPW, G>& S, int dir, int at, int Z, const LaneC<CPW>& L) {
  int v[4 * G];
  get_n<CPW, G>(S, L, v);
  if (dir > 0) gap_open<CPW, G>(v, at, Z, L);
  else gap_close<CPW, G>(v, at, Z, L);
#pragma unroll
  for (int g = 0; g < G; ++g) S.put4i(kStN, g, L.pos0, v[4 * g], v[4 * g + 1], v[4 * g + 2], v[4 * g + 3]);
}

// Flow: explicit Euler affine step Q <- Q + dt*kappa*Q (common/ensemble_kernel.cu:105-110)
template <int CPW, int G>
__device__ __forceinline__ void deform(Strands<G>& C, const EnsParams* P, float dt, int Z, const LaneC<CPW>& L) {
  float dk[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) dk[i] = dt * P->kappa[i];
#pragma unroll
  for (int j = 0; j < 4 * G; ++j) {
    if (L.pos(j) < Z) {
      const float x = C.qx[j], y = C.qy[j], z = C.qz[j];
      C.qx[j] = __builtin_fmaf(dk[2], z, __builtin_fmaf(dk[1], y, __builtin_fmaf(dk[0], x, x)));
      C.qy[j] = __builtin_fmaf(dk[5], z, __builtin_fmaf(dk[4], y, __builtin_fmaf(dk[3], x, y)));
      C.qz[j] = __builtin_fmaf(dk[8], z, __builtin_fmaf(dk[7], y, __builtin_fmaf(dk[6], x, z)));
    }
  }
}

// EQ sample (sxy, syz, sxz) = -sum 3 Qa Qb / N  (common/eq_ensemble_kernel.cu:281-296)
template <int CPW, int G, bool LT>
__device__ __forceinline__ void eq_sample(const Strands<G>& C, const Stage<CPW, G>& S, const Tabs<LT>& T, int Z,
                                          const LaneC<CPW>& L, float out[3]) {
  int n[4 * G];
  get_n<CPW, G>(S, L, n);
  float a[4 * G], b[4 * G], d[4 * G];
#pragma unroll
  for (int j = 0; j < 4 * G; ++j) {
    const bool on = L.pos(j) < Z;
    const float t3 = on ? T.T3(n[j]) : 0.0f;
    a[j] = on ? (C.qx[j] * C.qy[j]) * t3 : 0.0f;
    b[j] = on ? (C.qy[j] * C.qz[j]) * t3 : 0.0f;
    d[j] = on ? (C.qx[j] * C.qz[j]) * t3 : 0.0f;
  }
  out[0] = -chain_total<CPW, G>(a, Z, L);
  out[1] = -chain_total<CPW, G>(b, Z, L);
  out[2] = -chain_total<CPW, G>(d, Z, L);
}

// flow sample: xx yy zz xy yz zx Lpp Z (common/ensemble_kernel.cu:504-538)
template <int CPW, int G, bool LT>
__device__ __forceinline__ void flow_sample(const Strands<G>& C, const Stage<CPW, G>& S, const Tabs<LT>& T, int Z,
                                            const LaneC<CPW>& L, float out[8]) {
  int n[4 * G];
  get_n<CPW, G>(S, L, n);
#pragma unroll
  for (int comp = 0; comp < 7; ++comp) {
    float v[4 * G];
#pragma unroll
    for (int j = 0; j < 4 * G; ++j) {
      const bool on = L.pos(j) < Z;
      const float x = C.qx[j], y = C.qy[j], z = C.qz[j];
      const float t3 = on ? T.T3(n[j]) : 0.0f;
      float val;
      switch (comp) {
        case 0: val = (x * x) * t3; break;
        case 1: val = (y * y) * t3; break;
        case 2: val = (z * z) * t3; break;
        case 3: val = (x * y) * t3; break;
        case 4: val = (y * z) * t3; break;
        case 5: val = (x * z) * t3; break;
        default: val = sqrtf(q2_of(x, y, z)); break;
      }
      v[j] = on ? val : 0.0f;
    }
    const float tot = chain_total<CPW, G>(v, Z, L);
    out[comp] = comp < 6 ? -tot : tot;
  }
  out[7] = (float)Z;
}

// MSD sample: centre of mass relative to R1 (common/eq_ensemble_kernel.cu:297-323),
// sum_i (P_i + Q_i/2) N_i / Nk with P_i the canonical inclusive scan of Q at i-1;
// the caller adds R1.
template <int CPW, int G>
__device__ __forceinline__ void com_sample(const Strands<G>& C, const Stage<CPW, G>& S, int Z, float nkf,
                                           const LaneC<CPW>& L, float out[3]) {
  int n[4 * G];
  get_n<CPW, G>(S, L, n);
#pragma unroll
  for (int comp = 0; comp < 3; ++comp) {
    float q[4 * G];
#pragma unroll
    for (int j = 0; j < 4 * G; ++j) {
      const float x = comp == 0 ? C.qx[j] : comp == 1 ? C.qy[j] : C.qz[j];
      q[j] = L.pos(j) < Z ? x : 0.0f;
    }
    float I[4 * G], P[4 * G], v[4 * G];
    chain_scan<CPW, G>(q, I, Z, L);
    shift_prev<CPW, G>(I, P, L);
#pragma unroll
    for (int j = 0; j < 4 * G; ++j)
      v[j] = L.pos(j) < Z ? ((P[j] + q[j] * 0.5f) * (float)n[j]) / nkf : 0.0f;
    out[comp] = chain_total<CPW, G>(v, Z, L);
  }
}

// Record every pending sample of every chain of the wave: G(t) sample k while
// t > k*delta (the state after the first jump past k*delta); flow output k while
// t >= k*delta.
template <int CPW, int G, bool LT, bool EXT>
__device__ __forceinline__ void record_samples(const Strands<G>& C, const Stage<CPW, G>& S, const Tabs<LT>& T,
                                               const EnsLocal& E, Hdr& H, float* stage, const float* r1,
                                               const LaneC<CPW>& L) {
  if (E.sample_mode == kSampleEq || (EXT && E.sample_mode == kSampleMsd)) {
    const bool msd = EXT && E.sample_mode == kSampleMsd;
    for (;;) {
      const bool pend = !H.inert && H.t > H.t_next;
      if (!__ballot(pend)) break;
      float sv[3];
      if constexpr (EXT) {
        if (msd) com_sample<CPW, G>(C, S, H.Z, (float)H.nk, L, sv);
        else eq_sample<CPW, G, LT>(C, S, T, H.Z, L, sv);
      } else {
        eq_sample<CPW, G, LT>(C, S, T, H.Z, L, sv);
      }
      if (pend) {
        if (L.sl == 0 && H.next_sample >= E.k0) {
          float* o = stage + E.stage_off + ((H.next_sample - E.k0) * E.nchains + H.local) * 3;
          if (EXT && msd) {
            const float* r = r1 + 3 * (int64_t)H.c;
            o[0] = sv[0] + r[0]; o[1] = sv[1] + r[1]; o[2] = sv[2] + r[2];
          } else {
            o[0] = sv[0]; o[1] = sv[1]; o[2] = sv[2];
          }
        }
        ++H.next_sample;
        H.t_next = next_time(E, H.next_sample);
      }
    }
  } else if (E.sample_mode == kSampleFlow) {
    for (;;) {
      const bool pend = !H.inert && H.t >= H.t_next;
      if (!__ballot(pend)) break;
      float sv[8];
      flow_sample<CPW, G, LT>(C, S, T, H.Z, L, sv);
      if (pend) {
        if (L.sl == 0 && H.next_sample >= E.k0) {
          float* o = stage + E.stage_off + ((H.next_sample - E.k0) * E.nchains + H.local) * 8;
#pragma unroll
          for (int q = 0; q < 8; ++q) o[q] = sv[q];
        }
        ++H.next_sample;
        H.t_next = next_time(E, H.next_sample);
      }
    }
  }
}

// ------------------------------------------------------------------ one jump of every active chain
template <int CPW, int G, bool LT, bool EXT>
__device__ __forceinline__ void jump(Strands<G>& C, Hdr& H, bool act, const Stage<CPW, G>& S, const Tabs<LT>& T,
                                     const EnsLocal& E, RandCache& RC, int& k_rng, float* r1, const float* gtabs,
                                     const LaneC<CPW>& L) {
  constexpr int SEG = Geo<CPW, G>::SEG;
  constexpr int CAP = Geo<CPW, G>::CAP;
  const int Z = H.Z;
  if (k_rng == SEG) {
    refill_rand(RC, H.ev, L.sl, H.cid, E);
    k_rng = 0;
  }
  // 1. rates at positions: the rate cache holds the pair rates at 0..Z-2 and (pair and
  // wave forms) the sum of the chain-end rates at Z-1; the row form puts that sum there
  float v[4 * G];
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const vf4 x = S.get4(kStRR, g, L.pos0);
    v[4 * g] = x.x; v[4 * g + 1] = x.y; v[4 * g + 2] = x.z; v[4 * g + 3] = x.w;
  }
  EndRates er_row{0.0f, 0.0f, 0.0f, 0.0f, 0.0f};
  if constexpr (!kEndCache<CPW>) {
    er_row = end_rates<CPW, G, LT>(S, T, Z, H.e1);
    const float Eend = er_row.total();
#pragma unroll
    for (int j = 0; j < 4 * G; ++j) v[j] = (L.pos(j) == Z - 1) ? Eend : v[j];
  }
  // 2. scan, total, clock
  float Gs[4 * G];
  const float W = chain_scan<CPW, G>(v, Gs, Z, L);
  if constexpr (kStageG<CPW, G>) {
#pragma unroll
    for (int g = 0; g < G; ++g) S.put4(kStG, g, L.pos0, Gs[4 * g], Gs[4 * g + 1], Gs[4 * g + 2], Gs[4 * g + 3]);
  }
  const float u_sel = seg_get<CPW>(RC.sel, k_rng, L);
  float dt;
  if (E.exp_wait) {
    const float u_wait = seg_get<CPW>(RC.wait, k_rng, L);
    dt = (-dlogf(u_wait)) / W;
  } else {
    dt = 1.0f / W;
  }
  const float pr = W * u_sel;
  const int p = pick_position<CPW, G>(Gs, pr, Z, L);
  if constexpr (kStageG<CPW, G>) wave_lds_sync();
  // 3. sub-event (per lane, replicated per chain); the pair and wave forms compute the
  // chain-end rates one by one only where a chain picked its end position
  EndRates er = er_row;
  if constexpr (kEndCache<CPW>) {
    if (__ballot(act && p == Z - 1)) er = end_rates<CPW, G, LT>(S, T, Z, H.e1);
  }
  const float cdcz = er.cdcz, sdc1 = er.sdc1, sdcz = er.sdcz, sdd1 = er.sdd1, sddz = er.sddz;
  const int pa = p < CAP - 2 ? p : CAP - 2;  // in-range position for the reads below
  // the inclusive sum before pa: staged (row form), or its lane picks the slot and one
  // cross-lane read fetches it
  float Gprev;
  if constexpr (kStageG<CPW, G>) {
    Gprev = S.getf(kStG, pa > 0 ? pa - 1 : 0);
  } else {
    const int q = pa > 0 ? pa - 1 : 0;
    const int jq = 4 * (q / Geo<CPW, G>::GS) + (q & 3);
    float mine = Gs[0];
#pragma unroll
    for (int k = 1; k < 4 * G; ++k) mine = (jq == k) ? Gs[k] : mine;
    Gprev = seg_get<CPW>(mine, (q >> 2) & (SEG - 1), L);
  }
  const int Np = S.geti(kStN, pa), Np1 = S.geti(kStN, pa + 1);
  const float wpp = S.getf(kStWP, pa), wmp = S.getf(kStWM, pa);
  const float wcp = S.getf(kStW, pa), wcp1 = S.getf(kStW, pa + 1);
  int kind = kEvNone;
  float r = 0.0f;
  if (act) {
    int st = 0;
    if (dt == 0.0f) st = 1;
    if (dt != dt) st = 2;
    if (__float_as_uint(dt) == 0x7f800000u) st = 3;
    if (st) {
      H.stall = st;
      kind = kEvStall;
    } else {
      H.t += (double)dt;
      H.dt_prev = dt;
      H.ev += 1;
      H.zsum += (uint64_t)Z;
      ++H.done;
      --H.left;
      if (p < Z) {
        r = (p == 0) ? pr : pr - Gprev;
        const bool pair = p < Z - 1;
        const float s0 = pair ? (E.cd ? wcp : 0.0f) : cdcz;
        const float s1 = pair ? wpp : sdc1;
        const float s2 = pair ? wmp : sdcz;
        const float s3 = pair ? T.CC(Np) : sdd1;
        const float s4 = pair ? 0.0f : sddz;
        // first m with s[m] > 0 and the running remainder below s[m]
        int m = -1;
        float rr = r;
        if (s0 > 0.0f && rr < s0) { m = 0; } else { rr = rr - s0;
        if (s1 > 0.0f && rr < s1) { m = 1; } else { rr = rr - s1;
        if (s2 > 0.0f && rr < s2) { m = 2; } else { rr = rr - s2;
        if (s3 > 0.0f && rr < s3) { m = 3; } else { rr = rr - s3;
        if (s4 > 0.0f && rr < s4) { m = 4; } else { rr = rr - s4; } } } } }
        r = rr;
        if (pair) {
          kind = m == 0 ? kEvCDd : m == 1 ? kEvShiftP : m == 2 ? kEvShiftM : m == 3 ? kEvCDc : kEvNone;
        } else {
          kind = m == 0 ? kEvCDcTail : m == 1 ? kEvSDcHead : m == 2 ? kEvSDcTail : m == 3 ? kEvSDdHead
               : m == 4 ? kEvSDdTail : kEvNone;
        }
      }
    }
  }
  ++k_rng;
  // MSD: destroying the first entanglement moves R1 forward by the old Q_1
  // (common/eq_ensemble_kernel.cu:447-455,645-653)
  if (EXT && E.sample_mode == kSampleMsd) {
    const bool fwd = (kind == kEvCDd && p == 0) || kind == kEvSDdHead;
    if (__ballot(fwd)) {
      const float x = value_at<CPW, G>(C.qx, 1, L), y = value_at<CPW, G>(C.qy, 1, L);
      const float z = value_at<CPW, G>(C.qz, 1, L);
      if (fwd && L.sl == 0) {
        float* r = r1 + 3 * (int64_t)H.c;
        r[0] = r[0] + x; r[1] = r[1] + y; r[2] = r[2] + z;
      }
    }
  }
  // 4. effect.  Common case: an N shift across pair p (p <= Z - 2 < CAP - 2, so Np and
  // Np1 above are N at p and p + 1): the chain's first lane writes the two new lengths.
  int lo = 1, hi = 0, zero_at = -1;
  // the chain-end rates change with N0, N1, N[Z-2], N[Z-1] and Z
  bool ends = kind >= kEvCDd && kind <= kEvSDdTail;
  if (kind == kEvShiftP || kind == kEvShiftM) {
    ends = p <= 1 || p >= Z - 3;
    const int dn = (kind == kEvShiftP) ? 1 : -1;
    if (L.sl == 0) {
      S.puti(kStN, p, Np + dn);
      S.puti(kStN, p + 1, Np1 - dn);
    }
    lo = p - 1; hi = p + 1;
    H.dlo = min(H.dlo, p); H.dhi = max(H.dhi, p + 1);
  }
  const bool rare = kind >= kEvCDd && kind <= kEvSDdTail;
  if (__ballot(rare)) {
    // values the rare effects read (cross-lane: run on the whole wave).  The EXT
    // variants read each group only when a chain of the wave needs it (the strand
    // vectors at p, p+1 for CD events, the random numbers for creations): that keeps
    // them within three waves per SIMD; the lean ones read them unconditionally, which
    // measured slightly faster.
    const int pq = p < CAP - 1 ? p : CAP - 2;
    float qxp = 0.0f, qyp = 0.0f, qzp = 0.0f, qxp1 = 0.0f, qyp1 = 0.0f, qzp1 = 0.0f;
    if (!EXT || __ballot(kind == kEvCDd || kind == kEvCDc)) {
      qxp = value_at<CPW, G>(C.qx, pq, L); qyp = value_at<CPW, G>(C.qy, pq, L);
      qzp = value_at<CPW, G>(C.qz, pq, L);
      qxp1 = value_at<CPW, G>(C.qx, pq + 1, L); qyp1 = value_at<CPW, G>(C.qy, pq + 1, L);
      qzp1 = value_at<CPW, G>(C.qz, pq + 1, L);
    }
    float u_tau = 0.0f, g0 = 0.0f, g1 = 0.0f, g2 = 0.0f;
    if (!EXT || __ballot(kind == kEvCDc || kind == kEvCDcTail || kind == kEvSDcHead || kind == kEvSDcTail)) {
      u_tau = seg_get<CPW>(RC.tau, k_rng - 1, L);
      g0 = seg_get<CPW>(RC.g0, k_rng - 1, L); g1 = seg_get<CPW>(RC.g1, k_rng - 1, L);
      g2 = seg_get<CPW>(RC.g2, k_rng - 1, L);
    }
    const int N0n = S.geti(kStN, 0), N1n = S.geti(kStN, 1);
    const int Nz = S.geti(kStN, Z - 1), Nz2 = S.geti(kStN, Z >= 2 ? Z - 2 : 0);
    // polydisperse background: the new link's partner quantile is the top bits of word 3
    // of this jump's Philox draw (H.ev already counts the jump)
    int pidx = 0;
    const float* pd = gtabs + E.P->pd_off;
    const bool partner = EXT && E.partner;
    // (only creations draw a partner; the draw runs when a chain of the wave creates)
    if (partner && __ballot(kind == kEvCDc || kind == kEvCDcTail || kind == kEvSDcHead || kind == kEvSDcTail)) {
      const uint64_t evj = H.ev - 1;
      const uint4 x = philox4x32_10(make_uint4((uint32_t)evj, (uint32_t)(evj >> 32), H.cid, kPurposeStep), E.key0,
                                    E.key1);
      pidx = (int)(x.w >> (32 - kPartnerBits));
    }
    int gap = 0, at = CAP, znew = Z;
    int p0 = -1, p0N = 0, p1 = -1, p1N = 0, w0i = -1, w1i = -1;
    float p0x = 0.0f, p0y = 0.0f, p0z = 0.0f, p1x = 0.0f, p1y = 0.0f, p1z = 0.0f, w0 = 0.0f, w1 = 0.0f;
    if (rare) {
      const int e = p;
      if (kind == kEvCDd) {            // CD destruction of link e: merge strands e, e+1
        const int Ne = Np, Ne1 = Np1;
        float mx = 0.0f, my = 0.0f, mz = 0.0f;
        if (!(e == 0 || e == Z - 2)) { mx = qxp + qxp1; my = qyp + qyp1; mz = qzp + qzp1; }
        gap = -1; at = e + 1; znew = Z - 1;
        p0 = e; p0N = Ne + Ne1; p0x = mx; p0y = my; p0z = mz;
        w0i = e; w0 = wcp1;
        lo = e - 1; hi = e;
        // the old end position, past the end now (with the cached chain-end rates: the gap
        // close leaves them there), or the new one (the gap close shifts nothing when
        // e = Z-2)
        zero_at = kEndCache<CPW> ? Z - 1 : Z - 2;
        H.dlo = min(H.dlo, e); H.dhi = max(H.dhi, Z - 1);
      } else if (kind == kEvCDc) {     // CD creation inside strand e
        const int Ne = Np;
        const float wc = T.CC(Ne);
        const int newn = split_newn(r, Ne, wc);
        const float wnew = partner ? rate_fdt_partner(pd, pidx, u_tau) : rate_fdt(E.P, u_tau);
        gap = 1; at = e; znew = Z + 1;
        w0i = e; w0 = wnew;
        if (e == 0) {
          const float sig = sqrtf((float)(Ne - newn) / 3.0f);
          p0 = 0; p0N = newn;
          p1 = 1; p1N = Ne - newn; p1x = g0 * sig; p1y = g1 * sig; p1z = g2 * sig;
          lo = 0; hi = 1;
        } else {
          const float sig = sqrtf(((float)newn * (float)(Ne - newn)) / (3.0f * (float)Ne));
          const float ratio = (float)newn / (float)Ne;
          const float nx = (g0 * sig) + (qxp * ratio);
          const float ny = (g1 * sig) + (qyp * ratio);
          const float nz = (g2 * sig) + (qzp * ratio);
          p0 = e; p0N = newn; p0x = nx; p0y = ny; p0z = nz;
          p1 = e + 1; p1N = Ne - newn; p1x = qxp - nx; p1y = qyp - ny; p1z = qzp - nz;
          lo = e - 1; hi = e + 1;
        }
        H.dlo = min(H.dlo, e); H.dhi = max(H.dhi, Z);
      } else if (kind == kEvCDcTail) { // CD creation in the tail strand
        const int N = Nz;
        const int newn = split_newn(r, N, cdcz);
        const float sig = (Z == 1) ? 0.0f : sqrtf((float)newn / 3.0f);
        znew = Z + 1;
        p0 = Z - 1; p0N = newn;
        if (Z != 1) { p0x = g0 * sig; p0y = g1 * sig; p0z = g2 * sig; }
        w0i = Z - 1; w0 = partner ? rate_fdt_partner(pd, pidx, u_tau) : rate_fdt(E.P, u_tau);
        p1 = Z; p1N = N - newn;
        w1i = Z; w1 = 0.0f;
        lo = Z - 2; hi = Z - 1;
        H.dlo = min(H.dlo, Z - 1); H.dhi = max(H.dhi, Z);
      } else if (kind == kEvSDcHead) { // SD creation at the head
        const float wnew = partner ? rate_ft_partner(pd, pidx, u_tau) : rate_ft(E.P, u_tau);
        const float sig = (Z == 1) ? 0.0f : sqrtf((float)(N0n - 1) / 3.0f);
        gap = 1; at = 0; znew = Z + 1;
        p0 = 0; p0N = 1;
        w0i = 0; w0 = E.cd ? wnew : 0.0f;
        p1 = 1; p1N = N0n - 1;
        if (Z != 1) { p1x = g0 * sig; p1y = g1 * sig; p1z = g2 * sig; }
        lo = 0; hi = 1;
        H.dlo = min(H.dlo, 0); H.dhi = max(H.dhi, Z);
      } else if (kind == kEvSDcTail) { // SD creation at the tail
        const float wnew = partner ? rate_ft_partner(pd, pidx, u_tau) : rate_ft(E.P, u_tau);
        const float sig = (Z == 1) ? 0.0f : sqrtf((float)(Nz - 1) / 3.0f);
        znew = Z + 1;
        p0 = Z - 1; p0N = Nz - 1;
        if (Z != 1) { p0x = g0 * sig; p0y = g1 * sig; p0z = g2 * sig; }
        w0i = Z - 1; w0 = E.cd ? wnew : 0.0f;
        p1 = Z; p1N = 1;
        w1i = Z; w1 = 0.0f;
        lo = Z - 2; hi = Z - 1;
        H.dlo = min(H.dlo, Z - 1); H.dhi = max(H.dhi, Z);
      } else if (kind == kEvSDdHead) { // SD destruction at the head: strands 0,1 merge into an end
        gap = -1; at = 0; znew = Z - 1;
        p0 = 0; p0N = N1n + 1;
        lo = 0; hi = 0;
        if (kEndCache<CPW>) zero_at = Z - 1;  // the gap close leaves the cached end rates there
        H.dlo = min(H.dlo, 0); H.dhi = max(H.dhi, Z - 1);
      } else {                          // SD destruction at the tail
        znew = Z - 1;
        p0 = Z - 2; p0N = Nz2 + 1;
        w0i = Z - 2; w0 = 0.0f;
        lo = Z - 3; hi = Z - 3;
        zero_at = kEndCache<CPW> ? Z - 1 : Z - 2;  // the old (cached end rates) or new end position
        H.dlo = min(H.dlo, Z - 2); H.dhi = max(H.dhi, Z - 2);
      }
    }
    // gaps (whole wave; identity on chains without one).  The lengths move in LDS, where
    // the other chains of the wave may have just written a shift: order those writes first.
    wave_lds_sync();
    const bool any_open = __ballot(gap > 0) != 0, any_close = __ballot(gap < 0) != 0;
    if (any_open) {
      const int ao = gap > 0 ? at : CAP;
      gap_open<CPW, G>(C.qx, ao, Z, L); gap_open<CPW, G>(C.qy, ao, Z, L); gap_open<CPW, G>(C.qz, ao, Z, L);
      gap_open<CPW, G>(C.w, ao, Z, L);
      gap_n<CPW, G>(S, 1, ao, Z, L);
      if (!EXT || !E.flow) gap_lds<CPW, G>(S, 1, ao, Z, L);
    }
    if (any_close) {
      const int ac = gap < 0 ? at : CAP;
      gap_close<CPW, G>(C.qx, ac, Z, L); gap_close<CPW, G>(C.qy, ac, Z, L); gap_close<CPW, G>(C.qz, ac, Z, L);
      gap_close<CPW, G>(C.w, ac, Z, L);
      gap_n<CPW, G>(S, -1, ac, Z, L);
      if (!EXT || !E.flow) gap_lds<CPW, G>(S, -1, ac, Z, L);
    }
    // patches (the lengths by the chain's first lane, after the gap ops' LDS writes)
#pragma unroll
    for (int j = 0; j < 4 * G; ++j) {
      const int ps = L.pos(j);
      if (ps == p0) { C.qx[j] = p0x; C.qy[j] = p0y; C.qz[j] = p0z; }
      if (ps == p1) { C.qx[j] = p1x; C.qy[j] = p1y; C.qz[j] = p1z; }
      if (ps == w0i) C.w[j] = w0;
      if (ps == w1i) C.w[j] = w1;
    }
    if (L.sl == 0) {
      if (p0 >= 0) S.puti(kStN, p0, p0N);
      if (p1 >= 0) S.puti(kStN, p1, p1N);
    }
    // MSD: a first entanglement created before the old one moves R1 back by the new Q_1
    // (common/eq_ensemble_kernel.cu:531-539,606-614)
    if (EXT && E.sample_mode == kSampleMsd && L.sl == 0 && ((kind == kEvCDc && p == 0) || kind == kEvSDcHead)) {
      float* r = r1 + 3 * (int64_t)H.c;
      r[0] = r[0] - p1x; r[1] = r[1] - p1y; r[2] = r[2] - p1z;
    }
    H.Z = znew;
    stage_qw<CPW, G>(S, C, L);
  }
  // 5. refresh the pairs next to the change.  Under flow every pair rate is recomputed
  // from the deformed strands before the next jump (the affine step follows every jump),
  // so the EXT variant, which the passes with a flow ensemble launch, neither shifts the
  // cached rates (gap_lds above) nor patches them here; the lean variant keeps the
  // unconditional code, measured 1.3% faster for equilibrium passes.
  wave_lds_sync();
  if (!EXT || !E.flow) {
    refresh_pairs<CPW, G, LT>(S, T, E.cd, H.Z, lo, hi, zero_at, L);
    // the chain-end rates at the (new) end position, after refresh_pairs' writes
    if constexpr (kEndCache<CPW>) {
      if (__ballot(ends)) {
        const float eend = end_rates<CPW, G, LT>(S, T, H.Z, H.e1).total();
        if (ends && L.sl == 0) S.putf(kStRR, H.Z - 1, eend);
      }
    }
    wave_lds_sync();
  }
}

// ------------------------------------------------------------------ load / store
template <int CPW, int G>
__device__ __forceinline__ void load_strands(Strands<G>& C, const Stage<CPW, G>& S, const ChainArrays& ch,
                                             const Hdr& H, int64_t base, const LaneC<CPW>& L) {
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int p0 = Geo<CPW, G>::GS * g + L.pos0;
    float4 x = make_float4(0.f, 0.f, 0.f, 0.f), y = x, z = x, w = x;
    int4 n = make_int4(0, 0, 0, 0);
    if (!H.inert && p0 < H.Z) {
      const int64_t o = base + p0;
      x = *(const float4*)(ch.Qx + o); y = *(const float4*)(ch.Qy + o); z = *(const float4*)(ch.Qz + o);
      w = *(const float4*)(ch.w + o); n = *(const int4*)(ch.N + o);
    }
    const float xs[4] = {x.x, x.y, x.z, x.w}, ys[4] = {y.x, y.y, y.z, y.w}, zs[4] = {z.x, z.y, z.z, z.w};
    const float ws[4] = {w.x, w.y, w.z, w.w};
    const int ns[4] = {n.x, n.y, n.z, n.w};
    int nv[4];
#pragma unroll
    for (int k = 0; k < 4; ++k) {
      const int j = 4 * g + k;
      const bool on = p0 + k < H.Z;
      C.qx[j] = on ? xs[k] : 0.0f; C.qy[j] = on ? ys[k] : 0.0f; C.qz[j] = on ? zs[k] : 0.0f;
      C.w[j] = on ? ws[k] : 0.0f; nv[k] = on ? ns[k] : 0;
    }
    S.put4i(kStN, g, L.pos0, nv[0], nv[1], nv[2], nv[3]);
  }
}

// positions lo..hi (and < Z) back to HBM; whole 4-slot groups when they lie inside
template <int CPW, int G>
__device__ __forceinline__ void store_strands(const Strands<G>& C, const Stage<CPW, G>& S, const ChainArrays& ch,
                                              const Hdr& H, int64_t base, int lo, int hi, const LaneC<CPW>& L) {
  hi = min(hi, H.Z - 1);
  int n[4 * G];
  get_n<CPW, G>(S, L, n);
#pragma unroll
  for (int g = 0; g < G; ++g) {
    const int p0 = Geo<CPW, G>::GS * g + L.pos0;
    const int64_t o = base + p0;
    if (p0 >= lo && p0 + 3 <= hi) {
      *(float4*)(ch.Qx + o) = make_float4(C.qx[4 * g], C.qx[4 * g + 1], C.qx[4 * g + 2], C.qx[4 * g + 3]);
      *(float4*)(ch.Qy + o) = make_float4(C.qy[4 * g], C.qy[4 * g + 1], C.qy[4 * g + 2], C.qy[4 * g + 3]);
      *(float4*)(ch.Qz + o) = make_float4(C.qz[4 * g], C.qz[4 * g + 1], C.qz[4 * g + 2], C.qz[4 * g + 3]);
      *(float4*)(ch.w + o) = make_float4(C.w[4 * g], C.w[4 * g + 1], C.w[4 * g + 2], C.w[4 * g + 3]);
      *(int4*)(ch.N + o) = make_int4(n[4 * g], n[4 * g + 1], n[4 * g + 2], n[4 * g + 3]);
    } else {
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int j = 4 * g + k;
        if (p0 + k >= lo && p0 + k <= hi) {
          ch.Qx[o + k] = C.qx[j]; ch.Qy[o + k] = C.qy[j]; ch.Qz[o + k] = C.qz[j];
          ch.w[o + k] = C.w[j]; ch.N[o + k] = n[j];
        }
      }
    }
  }
}

// ------------------------------------------------------------------ a unit of work
// Runs the chains of one work unit (four chains in row form, one in wave form) until
// each reaches the ensemble's t_end (or its jump target ev_stop, or one jump in the
// plain form), pushing chains that outgrow the geometry to the overflow list.
template <int CPW, int G, bool LT, bool PLAIN, bool EXT>
__device__ __forceinline__ void run_unit(const KmcArgs& a, int unit, int n_entries, lds_f* stg, lds_f* ltab,
                                         int& lds_ens) {
  constexpr int CAP = Geo<CPW, G>::CAP;
  constexpr int SEG = Geo<CPW, G>::SEG;
  const LaneC<CPW> L;
  const int c0 = __builtin_amdgcn_readfirstlane(a.work[CPW * unit]);
  const int ens = __builtin_amdgcn_readfirstlane(a.ch.ens[c0]);
  EnsLocal E;
  load_ens(E, a.ens, ens);
  Tabs<LT> T;
  T.n = E.nmax + 2;
  if constexpr (LT) {
    if (lds_ens != ens) {
      const int tsz = 9 * T.n + 4 * E.nmax + 4;
      const float* src = a.tabs + E.P->tab_off;
      for (int i = L.lane; i < tsz; i += 64) ltab[i] = src[i];
      wave_lds_sync();
      lds_ens = ens;
    }
    T.b = ltab;
  } else {
    T.b = (typename TabPtr<false>::type)(a.tabs + E.P->tab_off);
  }
  Hdr H;
  {
    const int ent = CPW * unit + L.lane / Geo<CPW, G>::SEG;
    H.c = CPW == 1 ? c0 : (ent < n_entries ? a.work[ent] : -1);
  }
  H.ovf = 0; H.done = 0; H.dlo = 1 << 30; H.dhi = -1;
  if (H.c >= 0) {
    const ChainArrays& ch = a.ch;
    const int c = H.c;
    H.Z = ch.Z[c]; H.nk = ch.nk[c]; H.stall = ch.stall[c]; H.local = ch.local[c]; H.cid = ch.cid[c];
    H.next_sample = ch.next_sample[c]; H.t = ch.t[c]; H.ev = ch.ev[c];
    H.zsum = ch.zsum[c]; H.dt_prev = ch.dt_prev[c];
    const uint64_t stop = a.use_ev_stop ? ch.ev_stop[c] : ~0ull;
    H.left = stop <= H.ev ? 0 : (stop - H.ev > 0x7fffffffull ? 0x7fffffff : (int)(stop - H.ev));
    H.inert = H.stall != 0 || H.Z > CAP;
    if (H.stall == 0 && H.Z > CAP) H.ovf = 1;
  } else {
    H.Z = 1; H.nk = 1; H.stall = 0; H.local = 0; H.cid = 0; H.next_sample = 0; H.t = 0.0;
    H.ev = 0; H.zsum = 0; H.dt_prev = 0.0f; H.left = 0; H.inert = 1;
  }
  const bool inert0 = H.inert != 0;
  if (H.inert) H.Z = 1;
  H.t_next = next_time(E, H.next_sample);
  H.e1 = 1.0f / (E.beta * (float)H.nk);
  Stage<CPW, G> S{stg, L.srow * G};
  Strands<G> C;
  load_strands<CPW, G>(C, S, a.ch, H, H.c >= 0 ? a.ch.base[H.c] : 0, L);
  {
    float q2[4 * G];
    stage_qw<CPW, G>(S, C, L, q2);
    refresh_all<CPW, G, LT>(S, C, q2, T, E.cd, H.Z, H.e1, H.inert == 0, L);
  }
  wave_lds_sync();
  RandCache RC{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int k_rng = SEG;
  for (;;) {
    if (E.flow) {
      const bool dfm = !H.inert && H.dt_prev != 0.0f;
      if (__ballot(dfm)) {
        if (dfm) deform<CPW, G>(C, E.P, H.dt_prev, H.Z, L);
        float q2[4 * G];
        stage_qw<CPW, G>(S, C, L, q2);
        refresh_all<CPW, G, LT>(S, C, q2, T, E.cd, H.Z, H.e1, H.inert == 0, L);
        wave_lds_sync();
        if (dfm) { H.dlo = 0; H.dhi = max(H.dhi, H.Z - 1); }
      }
      H.dt_prev = 0.0f;
    }
    record_samples<CPW, G, LT, EXT>(C, S, T, E, H, a.stage, a.ch.R1, L);
    bool act = !H.inert && H.stall == 0 && H.t < E.t_end && H.left > 0 && !(PLAIN && H.done >= a.plain_jumps);
    if (act && H.Z >= CAP) { H.ovf = 1; act = false; }
    H.inert = H.inert || (H.ovf != 0);
    if (!__ballot(act)) break;
    jump<CPW, G, LT, EXT>(C, H, act, S, T, E, RC, k_rng, a.ch.R1, a.tabs, L);
  }
  // write back every chain that was loaded live
  if (H.c >= 0 && !inert0) {
    const int64_t base = a.ch.base[H.c];
    if (PLAIN) store_strands<CPW, G>(C, S, a.ch, H, base, H.dlo, H.dhi, L);
    else store_strands<CPW, G>(C, S, a.ch, H, base, 0, H.Z - 1, L);
    if (L.sl == 0) {
      const int c = H.c;
      a.ch.Z[c] = H.Z; a.ch.stall[c] = H.stall; a.ch.next_sample[c] = H.next_sample; a.ch.t[c] = H.t;
      a.ch.ev[c] = H.ev; a.ch.zsum[c] = H.zsum; a.ch.dt_prev[c] = H.dt_prev;
    }
  }
  if (H.c >= 0 && H.ovf && L.sl == 0) a.overflow[atomicAdd(a.n_overflow, 1)] = H.c;
}

// Batched form with per-segment refill.  Each segment of the wave (a 16-lane row in the
// row form, the whole wave in the wave form) takes chains one at a time from the work
// list: when its chain reaches the window end (or overflows the geometry) the chain is
// written back and the segment takes the next one, while the other segments keep
// jumping.  A quad of chains thus no longer waits for its slowest member.  A wave runs
// chains of one ensemble at a time: a segment that takes a chain of another ensemble
// holds it until the wave's running chains have finished, then the wave switches.
// Every (re)load restages the whole wave and recomputes every pair rate from the staged
// strands, which reproduces the cached rates of the running chains exactly, and refills
// the random-number cache from each chain's own jump count (the cache holds the Philox
// outputs of jumps ev .. ev+SEG-1 of each chain), so a chain's trajectory does not
// depend on which chains share its wave.
template <int CPW, int G, bool LT, bool EXT>
__device__ __forceinline__ void run_stream(const KmcArgs& a, lds_f* stg, lds_f* ltab) {
  constexpr int CAP = Geo<CPW, G>::CAP;
  constexpr int SEG = Geo<CPW, G>::SEG;
  // A segment's state lives in H.c and H.inert (no extra registers):
  //   H.c == -1: empty;  H.c == -2: past the end of the list;
  //   H.c >= 0 and H.inert == kHeld: chain taken, not loaded (another ensemble);
  //   H.inert == 0: running;  1: overflowed while running (written back, handed on);
  //   kDeadAtLoad: stalled or over capacity at load (not written back).
  constexpr int kHeld = 2, kDeadAtLoad = 3;
  const LaneC<CPW> L;
  const int n_work = *a.n_work;
  int lds_ens = -1, cur = -1;
  EnsLocal E;
  Tabs<LT> T;
  Hdr H;
  H.c = -1; H.Z = 1; H.nk = 1; H.stall = 0; H.local = 0; H.cid = 0; H.next_sample = 0; H.t = 0.0;
  H.t_next = 0.0; H.ev = 0; H.zsum = 0; H.dt_prev = 0.0f; H.left = 0; H.inert = 1; H.e1 = 0.0f;
  H.ovf = 0; H.done = 0; H.dlo = 1 << 30; H.dhi = -1;
  Stage<CPW, G> S{stg, L.srow * G};
  Strands<G> C;
#pragma unroll
  for (int j = 0; j < 4 * G; ++j) { C.qx[j] = 0.f; C.qy[j] = 0.f; C.qz[j] = 0.f; C.w[j] = 0.f; }
#pragma unroll
  for (int g = 0; g < G; ++g) S.put4i(kStN, g, L.pos0, 0, 0, 0, 0);
  RandCache RC{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
  int k_rng = SEG;
  for (;;) {
    // 1. empty segments take entries until they hold a chain or the list is exhausted
    for (;;) {
      const bool need = H.c == -1;
      if (!__ballot(need)) break;
      int ent = 0;
      if (need && L.sl == 0) ent = atomicAdd(a.queue, 1);
      ent = seg_get<CPW>(ent, 0, L);
      if (need) {
        const int c = ent < n_work ? a.work[ent] : -2;
        if (c != -1) { H.c = c; H.inert = kHeld; }
      }
    }
    // 2. the ensemble: switch when nothing of the current one runs or waits
    const bool held = H.c >= 0 && H.inert == kHeld;
    const int pe = held ? a.ch.ens[H.c] : -1;
    if (!__ballot(H.c >= 0 && (H.inert != kHeld || pe == cur))) {
      const uint64_t hm = __ballot(held);
      if (!hm) break;  // every segment is past the end of the list
      cur = __builtin_amdgcn_readlane(pe, (int)__builtin_ctzll(hm));
      load_ens(E, a.ens, cur);
      T.n = E.nmax + 2;
      if constexpr (LT) {
        if (lds_ens != cur) {
          const int tsz = 9 * T.n + 4 * E.nmax + 4;
          const float* src = a.tabs + E.P->tab_off;
          for (int i = L.lane; i < tsz; i += 64) ltab[i] = src[i];
          wave_lds_sync();
          lds_ens = cur;
        }
        T.b = ltab;
      } else {
        T.b = (typename TabPtr<false>::type)(a.tabs + E.P->tab_off);
      }
    }
    // 3. load the held chains of the current ensemble, restage the wave
    const bool ld = held && pe == cur;
    if (__ballot(ld)) {
      if (ld) {
        const ChainArrays& ch = a.ch;
        const int c = H.c;
        H.Z = ch.Z[c]; H.nk = ch.nk[c]; H.stall = ch.stall[c]; H.local = ch.local[c]; H.cid = ch.cid[c];
        H.next_sample = ch.next_sample[c]; H.t = ch.t[c]; H.ev = ch.ev[c];
        H.zsum = ch.zsum[c]; H.dt_prev = ch.dt_prev[c];
        const uint64_t stop = a.use_ev_stop ? ch.ev_stop[c] : ~0ull;
        H.left = stop <= H.ev ? 0 : (stop - H.ev > 0x7fffffffull ? 0x7fffffff : (int)(stop - H.ev));
        const bool dead = H.stall != 0 || H.Z > CAP;
        H.ovf = (H.stall == 0 && H.Z > CAP) ? 1 : 0;
        H.inert = dead ? kDeadAtLoad : 0;
        if (dead) H.Z = 1;
        H.t_next = next_time(E, H.next_sample);
        H.e1 = 1.0f / (E.beta * (float)H.nk);
        load_strands<CPW, G>(C, S, a.ch, H, a.ch.base[c], L);
      }
      wave_lds_sync();
      {
        float q2[4 * G];
        stage_qw<CPW, G>(S, C, L, q2);
        refresh_all<CPW, G, LT>(S, C, q2, T, E.cd, H.Z, H.e1, H.inert == 0, L);
      }
      wave_lds_sync();
      // the cache is refilled at every chain's current jump count before the next jump
      // (clearing it here tells the compiler the old values are dead in this branch)
      RC = RandCache{0.f, 0.f, 0.f, 0.f, 0.f, 0.f};
      k_rng = SEG;
    }
    // 4. jumps until a loaded chain finishes
    bool fin = false;
    for (;;) {
      if (E.flow) {
        const bool dfm = !H.inert && H.dt_prev != 0.0f;
        if (__ballot(dfm)) {
          if (dfm) deform<CPW, G>(C, E.P, H.dt_prev, H.Z, L);
          float q2[4 * G];
          stage_qw<CPW, G>(S, C, L, q2);
          refresh_all<CPW, G, LT>(S, C, q2, T, E.cd, H.Z, H.e1, H.inert == 0, L);
          wave_lds_sync();
        }
        H.dt_prev = 0.0f;
      }
      record_samples<CPW, G, LT, EXT>(C, S, T, E, H, a.stage, a.ch.R1, L);
      bool act = !H.inert && H.stall == 0 && H.t < E.t_end && H.left > 0;
      if (act && H.Z >= CAP) { H.ovf = 1; act = false; }
      if (H.ovf && !H.inert) H.inert = 1;
      fin = H.c >= 0 && H.inert != kHeld && !act;
      if (__ballot(fin)) break;
      jump<CPW, G, LT, EXT>(C, H, act, S, T, E, RC, k_rng, a.ch.R1, a.tabs, L);
    }
    // 5. write back the finished chains and hand on the overflowed ones
    if (fin) {
      if (H.inert != kDeadAtLoad) {
        store_strands<CPW, G>(C, S, a.ch, H, a.ch.base[H.c], 0, H.Z - 1, L);
        if (L.sl == 0) {
          const int c = H.c;
          a.ch.Z[c] = H.Z; a.ch.stall[c] = H.stall; a.ch.next_sample[c] = H.next_sample; a.ch.t[c] = H.t;
          a.ch.ev[c] = H.ev; a.ch.zsum[c] = H.zsum; a.ch.dt_prev[c] = H.dt_prev;
        }
      }
      if (H.ovf && L.sl == 0) a.overflow[atomicAdd(a.n_overflow, 1)] = H.c;
      H.c = -1; H.inert = 1; H.ovf = 0;
    }
  }
}

// LDS floats per wave: staging arrays, then the table copy
template <int CPW, int G>
__host__ __device__ constexpr int stage_floats() { return (kStArrays + (kStageG<CPW, G> ? 1 : 0)) * Geo<CPW, G>::STG; }

// Persistent kernel: each wave takes work units from the queue until it is empty.
// PLAIN: a.plain_jumps jumps per chain, the state in HBM between launches (the
// per-launch form; one jump per launch is the reference's structure); otherwise run to
// the targets.  Both forms are held to the same occupancy targets: three waves per SIMD
// (168 VGPRs) up to G = 2 (the G = 2 wave form spills 12 bytes per lane there; with 36
// spilled it finished the 262,144-chain melt's window in 117-126 ms against 130-131 ms at
// two waves: its chains, each a serial run of about 2,000 jumps per window, need waves in
// flight), one for G >= 4 (AGPRs as well).
// EXT: the variant with the less common features, launched when an ensemble of the pass
// needs one of them (it runs every other ensemble as well): MSD(t) (R1 tracking,
// centre-of-mass samples), the polydisperse CD background (partner tables) and flow
// (no incremental rate refresh: the affine step recomputes every rate).
template <int CPW, int G, bool LT, bool PLAIN, bool EXT>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu((CPW == 1 && G >= 4) ? 1 : 3))) kmc_kernel(KmcArgs a) {
  extern __shared__ float lds_dyn[];
  lds_f* stg = (lds_f*)lds_dyn;
  lds_f* ltab = stg + stage_floats<CPW, G>();
  // G >= 4: few long chains, each a serial run of thousands of jumps per window that
  // sets the pass's critical path.  Their waves take the SIMD's issue slot ahead of the
  // co-resident row, pair and G <= 2 waves, which are throughput work and fill the
  // cycles these leave idle.
  if constexpr (CPW == 1 && G >= 4 && !PLAIN) __builtin_amdgcn_s_setprio(3);
  if constexpr (!PLAIN && CPW > 1) {
    run_stream<CPW, G, LT, EXT>(a, stg, ltab);
  } else {
    int lds_ens = -1;
    // the list length is read once: its counter shares a cache line with the queue
    // counter every wave increments (lists of the larger geometries are not padded)
    const int n_entries = *a.n_work;
    const int n_units = (n_entries + CPW - 1) / CPW;
    for (;;) {
      int unit = 0;
      if (lane_id() == 0) unit = atomicAdd(a.queue, 1);
      unit = __builtin_amdgcn_readfirstlane(unit);
      if (unit >= n_units) break;
      run_unit<CPW, G, LT, PLAIN, EXT>(a, unit, n_entries, stg, ltab, lds_ens);
    }
  }
}

// Bins 0 and 1 of a concurrent pass in one launch: each wave takes chains from the bin-0
// list (Z < 64) in the row form until it is exhausted, then from the bin-1 list
// (64 <= Z < 128).  ROW2 (one-ensemble passes): bin 1 in the row form with two register
// groups, four chains per wave (the 262,144-chain melt's window: 80-82 ms against
// 89.7-90.3 ms in the pair form); otherwise in the pair form, two chains per wave (a
// wave runs one ensemble at a time, and in a serving mix four segments wait for each
// other more than two: the C5 stream at horizon 60 ran at 7.9e8 jumps/s with the
// two-group row form, 8.4e8 with the pair form).  Both shapes of a launch run within
// the same register budget (three waves per SIMD; the two-group row shape spills about
// 116 bytes per lane) and the larger staging size (12.3 KB with it, within the LDS for
// three waves per SIMD), and the launch leaves
// the pass's other hardware queues to the wave-form geometries: the long-chain runs,
// which set a window's critical path, and the wave-form bulk then all start at once.
// Each wave restages and reloads its rate tables when it moves to the bin-1 list, so
// every chain runs exactly as in separate launches.
template <bool ROW2>
constexpr int row_pair_stage() {
  return ROW2 ? (stage_floats<4, 1>() > stage_floats<4, 2>() ? stage_floats<4, 1>() : stage_floats<4, 2>())
              : (stage_floats<4, 1>() > stage_floats<2, 1>() ? stage_floats<4, 1>() : stage_floats<2, 1>());
}
template <bool LT, bool EXT, bool ROW2>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(3))) kmc_rowpair_kernel(KmcArgs row,
                                                                                                 KmcArgs pair) {
  extern __shared__ float lds_dyn[];
  lds_f* stg = (lds_f*)lds_dyn;
  lds_f* ltab = stg + row_pair_stage<ROW2>();
  run_stream<4, 1, LT, EXT>(row, stg, ltab);
  if constexpr (ROW2) run_stream<4, 2, LT, EXT>(pair, stg, ltab);
  else run_stream<2, 1, LT, EXT>(pair, stg, ltab);
}

}  // namespace dsm
