// dsm_kmc.h -- the wave-per-chain KMC step, templated on S (register slots per lane).
// Included by dsm_kmc_inst.hip (one translation unit per S).
//
// One wave64 owns one chain.  Strand i sits in lane i%64 of register slot i/64, so a
// chain of up to 64*S strands lives in VGPRs.  A jump (ktaletsk/gpu_dsm
// common/eq_ensemble_kernel.cu:242-663, common/ensemble_kernel.cu:188-502) is
//   1. cached pair rates at positions 0..Z-2, the five chain-end rates summed at Z-1,
//   2. a wave64 inclusive scan (DPP tree) -> total W, dt = 1/W, FP64 clock += dt,
//   3. pr = W*u with u from Philox(seed, chain, jump); the pair is the first position
//      whose inclusive sum exceeds pr (ballot + ffs); the sub-event follows the
//      reference order CDd, SD+, SD-, CDc (pairs) / CDc tail, SDc head/tail, SDd head/tail,
//   4. the event is reduced to a small uniform plan (an N shift, or a one-lane gap
//      opened/closed in every strand array, plus up to two strand and two link
//      patches) that one shared code path applies, and
//   5. the cached rates of the pairs next to the change are recomputed.
#pragma once
#include "dsm_device.h"
#include "dsm_types.h"

namespace dsm {

// FP32 tables of one ensemble, indexed by the integer strand length N:
// SP=.75/(N(N+1)), SM=.75/(N(N-1)), PP=(N/(N+1))^.75, PM=(N/(N-1))^.75, EC, ED, ED2
// (end rates), T3=3/N, CC=c(N-1); then FR[x]=4/x, x=2(f_i+f_j).  Contiguous, so one
// base pointer (LDS copy or global) and the row length n = nmax+2 address all of them.
struct Tabs {
  const float* b;
  int n;
  __device__ __forceinline__ float SP(int N) const { return b[N]; }
  __device__ __forceinline__ float SM(int N) const { return b[n + N]; }
  __device__ __forceinline__ float PP(int N) const { return b[2 * n + N]; }
  __device__ __forceinline__ float PM(int N) const { return b[3 * n + N]; }
  __device__ __forceinline__ float EC(int N) const { return b[4 * n + N]; }
  __device__ __forceinline__ float ED(int N) const { return b[5 * n + N]; }
  __device__ __forceinline__ float ED2(int N) const { return b[6 * n + N]; }
  __device__ __forceinline__ float T3(int N) const { return b[7 * n + N]; }
  __device__ __forceinline__ float CC(int N) const { return b[8 * n + N]; }
  __device__ __forceinline__ float FR(int x) const { return b[9 * n + x]; }
};

// Per-ensemble values used on every jump.  Rarely used ones (kappa, CD lifetime
// sampler constants) are read from the EnsParams record where they are needed.
struct EnsLocal {
  const EnsParams* P;
  int cd, flow, exp_wait, sample_mode, nchains;
  uint32_t key0, key1;
  double delta, t_end;
  int64_t k0, k_end, stage_off;
};

__device__ __forceinline__ void load_ens(EnsLocal& E, const EnsParams* __restrict__ ps, int e) {
  const EnsParams& P = ps[e];
  E.P = &P;
  E.cd = P.cd; E.flow = P.flow; E.exp_wait = P.exp_wait; E.sample_mode = P.sample_mode;
  E.nchains = P.nchains; E.key0 = P.key0; E.key1 = P.key1; E.delta = P.delta; E.t_end = P.t_end;
  E.k0 = P.k0; E.k_end = P.k_end; E.stage_off = P.stage_off;
}

// CD lifetime samplers returning the rate 1/tau (common/gpu_random.cu:143-148)
__device__ __forceinline__ float rate_ft(const EnsParams* P, float u) {
  return u < P->s_one_minus_g ? dpowf(__builtin_fmaf(u, P->s_At, P->s_Dt), P->s_Ct) : P->s_taud_inv;
}
__device__ __forceinline__ float rate_fdt(const EnsParams* P, float u) {
  return u < P->s_Bdt ? dpowf(__builtin_fmaf(u, P->s_Adt, P->s_Ddt), P->s_Cdt) : P->s_taud_inv;
}

// Sliding-dynamics shift rates of the pair (i, j=i+1) and the pair's total with its
// CD terms (common/eq_ensemble_kernel.cu:97-126).
__device__ __forceinline__ void pair_rate(const Tabs& T, int cd, int Ni, int Nj, float q2i, float q2j,
                                          float we, bool endi, bool endj, float& wp, float& wm,
                                          float& R) {
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

template <int S>
struct Chain {
  float qx[S], qy[S], qz[S], q2[S], w[S], wp[S], wm[S], rr[S];
  int n[S];
  int Z;
};

// record at the (uniform) position pos
template <int S, typename T>
__device__ __forceinline__ T rd(const T (&a)[S], int pos) {
  const int s = pos >> 6;
  T v = a[0];
#pragma unroll
  for (int k = 1; k < S; ++k) v = (s == k) ? a[k] : v;
  return rdlane(v, pos & 63);
}

// new[i] = (i > at && i <= Z) ? old[i-1] : old[i]
template <int S, typename T>
__device__ __forceinline__ void gap_open_arr(T (&a)[S], int at, int Z) {
  const int lane = lane_id();
#pragma unroll
  for (int s = S - 1; s >= 0; --s) {
    const T fill = (s > 0) ? rdlane(a[s > 0 ? s - 1 : 0], 63) : (T)0;
    const T pv = from_prev_lane(a[s], fill);
    const int i = s * 64 + lane;
    a[s] = (i > at && i <= Z) ? pv : a[s];
  }
}
// new[i] = (i >= at && i < Z-1) ? old[i+1] : old[i]
template <int S, typename T>
__device__ __forceinline__ void gap_close_arr(T (&a)[S], int at, int Z) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const T fill = (s + 1 < S) ? rdlane(a[s + 1 < S ? s + 1 : s], 0) : (T)0;
    const T nv = from_next_lane(a[s], fill);
    const int i = s * 64 + lane;
    a[s] = (i >= at && i < Z - 1) ? nv : a[s];
  }
}

// recompute cached rates of pairs lo..hi (clamped to 0..Z-2)
template <int S>
__device__ __forceinline__ void refresh(Chain<S>& C, const Tabs& T, int cd, int lo, int hi) {
  if (lo < 0) lo = 0;
  if (hi > C.Z - 2) hi = C.Z - 2;
  if (lo > hi) return;
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (s * 64 > hi || s * 64 + 63 < lo) continue;
    const int nfill = (s + 1 < S) ? rdlane(C.n[s + 1 < S ? s + 1 : s], 0) : 0;
    const float qfill = (s + 1 < S) ? rdlane(C.q2[s + 1 < S ? s + 1 : s], 0) : 0.0f;
    const int nj = from_next_lane(C.n[s], nfill);
    const float q2j = from_next_lane(C.q2[s], qfill);
    const int i = s * 64 + lane;
    if (i >= lo && i <= hi) {
      float wp, wm, R;
      pair_rate(T, cd, C.n[s], nj, C.q2[s], q2j, C.w[s], i == 0, i + 1 == C.Z - 1, wp, wm, R);
      C.wp[s] = wp; C.wm[s] = wm; C.rr[s] = R;
    }
  }
}

// inclusive tree-sum value at position Z-1 of per-position values v (zero beyond Z)
template <int S>
__device__ __forceinline__ float tree_total(const float (&v)[S], int Z) {
  const int last = Z - 1, ls = last >> 6;
  float carry = 0.0f, res = 0.0f;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (s <= ls) {
      const float sc = wave_scan(v[s]);
      const float at = rdlane(sc, last & 63), top = rdlane(sc, 63);
      if (s == ls) res = (s == 0) ? at : carry + at;
      carry = (s == 0) ? top : carry + top;
    }
  }
  return res;
}

// chain-end rates (common/eq_ensemble_kernel.cu:340-392): cdcz sdc1 sdcz sdd1 sddz
template <int S>
__device__ __forceinline__ float end_rates(const Chain<S>& C, const Tabs& T, float e1, float er[5]) {
  const int Z = C.Z;
  float cdcz = 0.0f, sdc1 = 0.0f, sdcz = 0.0f, sdd1 = 0.0f, sddz = 0.0f;
  if (Z == 1) {
    cdcz = T.CC(rd(C.n, 0));
    sdc1 = e1;
    sdcz = e1;
  } else {
    const int N0 = rd(C.n, 0), N1 = rd(C.n, 1), Nz = rd(C.n, Z - 1), Nz2 = rd(C.n, Z - 2);
    if (N0 == 1) sdd1 = (Z == 2) ? T.ED2(N1) : T.ED(N1);
    else sdc1 = T.EC(N0);
    if (Nz == 1) sddz = (Z == 2) ? T.ED2(Nz2) : T.ED(Nz2);
    else sdcz = T.EC(Nz);
    cdcz = T.CC(Nz);
  }
  er[0] = cdcz; er[1] = sdc1; er[2] = sdcz; er[3] = sdd1; er[4] = sddz;
  return (((cdcz + sdc1) + sdcz) + sdd1) + sddz;
}

__device__ __forceinline__ int pick_sub(const float* s, int n, float& r) {
  float rr = r;
  for (int m = 0; m < n; ++m) {
    if (s[m] > 0.0f && rr < s[m]) { r = rr; return m; }
    rr = rr - s[m];
  }
  r = rr;
  return -1;
}

__device__ __forceinline__ int split_newn(float r, int N, float wc) {
  int newn = (int)floorf(0.5f + (r * (float)(N - 2)) / wc) + 1;
  newn = newn < 1 ? 1 : newn;
  newn = newn > N - 1 ? N - 1 : newn;
  return newn;
}

// Philox outputs for 64 consecutive jumps of the chain, one jump per lane.
struct RandCache {
  float sel, wait, tau, g0, g1, g2;
  uint64_t base;
  bool valid;
};

__device__ __forceinline__ void refill_rand(RandCache& R, uint64_t ev, uint32_t cid, const EnsLocal& E) {
  const uint64_t e = ev + (uint64_t)lane_id();
  const uint32_t lo = (uint32_t)e, hi = (uint32_t)(e >> 32);
  const uint4 x = philox4x32_10(make_uint4(lo, hi, cid, kPurposeStep), E.key0, E.key1);
  R.sel = u01(x.x); R.wait = u01(x.y); R.tau = u01(x.z);
  const uint4 g = philox4x32_10(make_uint4(lo, hi, cid, kPurposeGauss), E.key0, E.key1);
  float g3;
  boxmuller(g.x, g.y, R.g0, R.g1);
  boxmuller(g.z, g.w, R.g2, g3);
  R.base = ev;
  R.valid = true;
}

// Flow: explicit Euler affine step Q <- Q + dt*kappa*Q (common/ensemble_kernel.cu:105-110)
template <int S>
__device__ __forceinline__ void deform(Chain<S>& C, const EnsParams* P, float dt) {
  float dk[9];
#pragma unroll
  for (int i = 0; i < 9; ++i) dk[i] = dt * P->kappa[i];
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    if (s * 64 + lane < C.Z) {
      const float x = C.qx[s], y = C.qy[s], z = C.qz[s];
      C.qx[s] = __builtin_fmaf(dk[2], z, __builtin_fmaf(dk[1], y, __builtin_fmaf(dk[0], x, x)));
      C.qy[s] = __builtin_fmaf(dk[5], z, __builtin_fmaf(dk[4], y, __builtin_fmaf(dk[3], x, y)));
      C.qz[s] = __builtin_fmaf(dk[8], z, __builtin_fmaf(dk[7], y, __builtin_fmaf(dk[6], x, z)));
      C.q2[s] = q2_of(C.qx[s], C.qy[s], C.qz[s]);
    }
  }
}

// EQ sample (sxy, syz, sxz) = -sum 3 Qa Qb / N  (common/eq_ensemble_kernel.cu:281-296)
template <int S>
__device__ __forceinline__ void eq_sample(const Chain<S>& C, const Tabs& T, float out[3]) {
  float a[S], b[S], d[S];
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const bool on = s * 64 + lane < C.Z;
    const float t3 = on ? T.T3(C.n[s]) : 0.0f;
    a[s] = on ? (C.qx[s] * C.qy[s]) * t3 : 0.0f;
    b[s] = on ? (C.qy[s] * C.qz[s]) * t3 : 0.0f;
    d[s] = on ? (C.qx[s] * C.qz[s]) * t3 : 0.0f;
  }
  out[0] = -tree_total<S>(a, C.Z);
  out[1] = -tree_total<S>(b, C.Z);
  out[2] = -tree_total<S>(d, C.Z);
}

// flow sample: xx yy zz xy yz zx Lpp Z (common/ensemble_kernel.cu:504-538)
template <int S>
__device__ __forceinline__ void flow_sample(const Chain<S>& C, const Tabs& T, float out[8]) {
  const int lane = lane_id();
#pragma unroll
  for (int comp = 0; comp < 7; ++comp) {
    float v[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const bool on = s * 64 + lane < C.Z;
      const float x = C.qx[s], y = C.qy[s], z = C.qz[s];
      const float t3 = on ? T.T3(C.n[s]) : 0.0f;
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
      v[s] = on ? val : 0.0f;
    }
    const float tot = tree_total<S>(v, C.Z);
    out[comp] = comp < 6 ? -tot : tot;
  }
  out[7] = (float)C.Z;
}

// Record every pending sample: G(t) sample k while t > k*delta (the state after the
// first jump past k*delta); flow output k while t >= k*delta.
template <int S>
__device__ __forceinline__ void record_samples(const Chain<S>& C, const Tabs& T, const EnsLocal& E,
                                               double t, int64_t& next_sample, int local, float* stage) {
  if (E.sample_mode == kSampleEq) {
    while (next_sample < E.k_end && t > (double)next_sample * E.delta) {
      float sv[3];
      eq_sample<S>(C, T, sv);
      if (lane_id() == 0 && next_sample >= E.k0) {
        float* o = stage + E.stage_off + ((next_sample - E.k0) * E.nchains + local) * 3;
        o[0] = sv[0]; o[1] = sv[1]; o[2] = sv[2];
      }
      ++next_sample;
    }
  } else if (E.sample_mode == kSampleFlow) {
    while (next_sample < E.k_end && t >= (double)next_sample * E.delta) {
      float sv[8];
      flow_sample<S>(C, T, sv);
      if (lane_id() == 0 && next_sample >= E.k0) {
        float* o = stage + E.stage_off + ((next_sample - E.k0) * E.nchains + local) * 8;
#pragma unroll
        for (int q = 0; q < 8; ++q) o[q] = sv[q];
      }
      ++next_sample;
    }
  }
}

// A jump's effect as uniform values (see the header comment).
struct Plan {
  int kind;
  int gap;           // 0 none, +1 open at `at`, -1 close at `at`
  int at;
  int znew;
  int dn_idx, dn;    // N[dn_idx] += dn, N[dn_idx+1] -= dn
  int p0, p0N;       // strand patch 0 (p0 < 0: none)
  float p0x, p0y, p0z;
  int p1, p1N;       // strand patch 1
  float p1x, p1y, p1z;
  int w0i, w1i;      // link patches
  float w0, w1;
  int lo, hi;        // pairs to recompute
  int dlo, dhi;      // strand records to write back (plain form)
};

// Steps 1-3: pick the jump and describe its effect.  Advances t, ev, dt_prev.
template <int S>
__device__ __forceinline__ void kmc_pick(const Chain<S>& C, const Tabs& T, const EnsLocal& E, float e1,
                                         uint32_t cid, RandCache& RC, double& t, uint64_t& ev, float& dt_prev,
                                         int& stall, Plan& L) {
  const int lane = lane_id();
  const int Z = C.Z;
  L.kind = kEvNone; L.gap = 0; L.at = 0; L.znew = Z; L.dn = 0; L.dn_idx = 0;
  L.p0 = -1; L.p1 = -1; L.w0i = -1; L.w1i = -1; L.lo = 1; L.hi = 0; L.dlo = 1 << 30; L.dhi = -1;
  L.p0N = L.p1N = 0; L.p0x = L.p0y = L.p0z = L.p1x = L.p1y = L.p1z = L.w0 = L.w1 = 0.0f;
  float er[5];
  const float Eend = end_rates<S>(C, T, e1, er);
  float incl[S];
  {
    float carry = 0.0f;
    const int ls = (Z - 1) >> 6;
#pragma unroll
    for (int s = 0; s < S; ++s) {
      incl[s] = 0.0f;
      if (s <= ls) {
        const int i = s * 64 + lane;
        const float pos = (i < Z - 1) ? C.rr[s] : (i == Z - 1 ? Eend : 0.0f);
        const float sc = wave_scan(pos);
        incl[s] = (s == 0) ? sc : carry + sc;
        const float top = rdlane(sc, 63);
        carry = (s == 0) ? top : carry + top;
      }
    }
  }
  const float W = rd(incl, Z - 1);
  if (!RC.valid || ev - RC.base >= 64) refill_rand(RC, ev, cid, E);
  const int k = (int)(ev - RC.base);
  const float u_sel = rdlane(RC.sel, k);
  const float dt = E.exp_wait ? (-dlogf(rdlane(RC.wait, k))) / W : 1.0f / W;
  int st = 0;
  if (dt == 0.0f) st = 1;
  if (dt != dt) st = 2;
  if (__float_as_uint(dt) == 0x7f800000u) st = 3;
  if (st) { stall = st; L.kind = kEvStall; return; }
  t += (double)dt;
  dt_prev = dt;
  ev += 1;
  const float pr = W * u_sel;
  int p = -1;
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const unsigned long long m = __ballot(incl[s] > pr && (s * 64 + lane) < Z);
    if (p < 0 && m) p = s * 64 + (__ffsll((long long)m) - 1);
  }
  if (p < 0) return;
  float r = (p == 0) ? pr : pr - rd(incl, p - 1);
  const float u_tau = rdlane(RC.tau, k);
  const float g0 = rdlane(RC.g0, k), g1 = rdlane(RC.g1, k), g2 = rdlane(RC.g2, k);
  if (p < Z - 1) {
    const int e = p;
    const int Ne = rd(C.n, e);
    float sub[4];
    sub[0] = E.cd ? rd(C.w, e) : 0.0f;
    sub[1] = rd(C.wp, e);
    sub[2] = rd(C.wm, e);
    sub[3] = T.CC(Ne);
    const int m = pick_sub(sub, 4, r);
    if (m == 1 || m == 2) {          // SD shift
      L.kind = (m == 1) ? kEvShiftP : kEvShiftM;
      L.dn_idx = e; L.dn = (m == 1) ? 1 : -1;
      L.lo = e - 1; L.hi = e + 1; L.dlo = e; L.dhi = e + 1;
    } else if (m == 0) {             // CD destruction: merge strands e, e+1
      const int Nm = Ne + rd(C.n, e + 1);
      float mx = 0.0f, my = 0.0f, mz = 0.0f;
      if (!(e == 0 || e == Z - 2)) {
        mx = rd(C.qx, e) + rd(C.qx, e + 1);
        my = rd(C.qy, e) + rd(C.qy, e + 1);
        mz = rd(C.qz, e) + rd(C.qz, e + 1);
      }
      L.kind = kEvCDd; L.gap = -1; L.at = e + 1; L.znew = Z - 1;
      L.p0 = e; L.p0N = Nm; L.p0x = mx; L.p0y = my; L.p0z = mz;
      L.w0i = e; L.w0 = rd(C.w, e + 1);
      L.lo = e - 1; L.hi = e; L.dlo = e; L.dhi = Z - 1;
    } else if (m == 3) {             // CD creation inside strand e
      const int newn = split_newn(r, Ne, sub[3]);
      const float wnew = rate_fdt(E.P, u_tau);
      L.kind = kEvCDc; L.gap = 1; L.at = e; L.znew = Z + 1;
      L.w0i = e; L.w0 = wnew;
      if (e == 0) {
        const float sig = sqrtf((float)(Ne - newn) / 3.0f);
        L.p0 = 0; L.p0N = newn;
        L.p1 = 1; L.p1N = Ne - newn; L.p1x = g0 * sig; L.p1y = g1 * sig; L.p1z = g2 * sig;
        L.lo = 0; L.hi = 1;
      } else {
        const float sig = sqrtf(((float)newn * (float)(Ne - newn)) / (3.0f * (float)Ne));
        const float ratio = (float)newn / (float)Ne;
        const float qx = rd(C.qx, e), qy = rd(C.qy, e), qz = rd(C.qz, e);
        const float nx = (g0 * sig) + (qx * ratio);
        const float ny = (g1 * sig) + (qy * ratio);
        const float nz = (g2 * sig) + (qz * ratio);
        L.p0 = e; L.p0N = newn; L.p0x = nx; L.p0y = ny; L.p0z = nz;
        L.p1 = e + 1; L.p1N = Ne - newn; L.p1x = qx - nx; L.p1y = qy - ny; L.p1z = qz - nz;
        L.lo = e - 1; L.hi = e + 1;
      }
      L.dlo = e; L.dhi = Z;
    }
  } else {
    const int m = pick_sub(er, 5, r);
    if (m == 0) {                    // CD creation in the tail strand
      const int N = rd(C.n, Z - 1);
      const int newn = split_newn(r, N, er[0]);
      const float sig = (Z == 1) ? 0.0f : sqrtf((float)newn / 3.0f);
      L.kind = kEvCDcTail; L.znew = Z + 1;
      L.p0 = Z - 1; L.p0N = newn;
      if (Z != 1) { L.p0x = g0 * sig; L.p0y = g1 * sig; L.p0z = g2 * sig; }
      L.w0i = Z - 1; L.w0 = rate_fdt(E.P, u_tau);
      L.p1 = Z; L.p1N = N - newn;
      L.w1i = Z; L.w1 = 0.0f;
      L.lo = Z - 2; L.hi = Z - 1; L.dlo = Z - 1; L.dhi = Z;
    } else if (m == 1) {             // SD creation at the head
      const int N0 = rd(C.n, 0);
      const float wnew = rate_ft(E.P, u_tau);
      const float sig = (Z == 1) ? 0.0f : sqrtf((float)(N0 - 1) / 3.0f);
      L.kind = kEvSDcHead; L.gap = 1; L.at = 0; L.znew = Z + 1;
      L.p0 = 0; L.p0N = 1;
      L.w0i = 0; L.w0 = E.cd ? wnew : 0.0f;
      L.p1 = 1; L.p1N = N0 - 1;
      if (Z != 1) { L.p1x = g0 * sig; L.p1y = g1 * sig; L.p1z = g2 * sig; }
      L.lo = 0; L.hi = 1; L.dlo = 0; L.dhi = Z;
    } else if (m == 2) {             // SD creation at the tail
      const int Nz = rd(C.n, Z - 1);
      const float wnew = rate_ft(E.P, u_tau);
      const float sig = (Z == 1) ? 0.0f : sqrtf((float)(Nz - 1) / 3.0f);
      L.kind = kEvSDcTail; L.znew = Z + 1;
      L.p0 = Z - 1; L.p0N = Nz - 1;
      if (Z != 1) { L.p0x = g0 * sig; L.p0y = g1 * sig; L.p0z = g2 * sig; }
      L.w0i = Z - 1; L.w0 = E.cd ? wnew : 0.0f;
      L.p1 = Z; L.p1N = 1;
      L.w1i = Z; L.w1 = 0.0f;
      L.lo = Z - 2; L.hi = Z - 1; L.dlo = Z - 1; L.dhi = Z;
    } else if (m == 3) {             // SD destruction at the head: strands 0,1 merge into an end
      L.kind = kEvSDdHead; L.gap = -1; L.at = 0; L.znew = Z - 1;
      L.p0 = 0; L.p0N = rd(C.n, 1) + 1;
      L.lo = 0; L.hi = 0; L.dlo = 0; L.dhi = Z - 1;
    } else if (m == 4) {             // SD destruction at the tail
      L.kind = kEvSDdTail; L.znew = Z - 1;
      L.p0 = Z - 2; L.p0N = rd(C.n, Z - 2) + 1;
      L.w0i = Z - 2; L.w0 = 0.0f;
      L.lo = Z - 3; L.hi = Z - 3; L.dlo = Z - 2; L.dhi = Z - 2;
    }
  }
}

// Steps 4-5: apply a plan to the register-resident chain.
template <int S>
__device__ __forceinline__ void kmc_apply(Chain<S>& C, const Tabs& T, int cd, const Plan& L) {
  const int lane = lane_id();
  if (L.dn != 0) {
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int i = s * 64 + lane;
      C.n[s] += (i == L.dn_idx) ? L.dn : (i == L.dn_idx + 1 ? -L.dn : 0);
    }
  } else if (L.gap != 0) {
    if (L.gap > 0) {
      gap_open_arr<S>(C.qx, L.at, C.Z); gap_open_arr<S>(C.qy, L.at, C.Z); gap_open_arr<S>(C.qz, L.at, C.Z);
      gap_open_arr<S>(C.q2, L.at, C.Z); gap_open_arr<S>(C.w, L.at, C.Z); gap_open_arr<S>(C.wp, L.at, C.Z);
      gap_open_arr<S>(C.wm, L.at, C.Z); gap_open_arr<S>(C.rr, L.at, C.Z); gap_open_arr<S>(C.n, L.at, C.Z);
    } else {
      gap_close_arr<S>(C.qx, L.at, C.Z); gap_close_arr<S>(C.qy, L.at, C.Z); gap_close_arr<S>(C.qz, L.at, C.Z);
      gap_close_arr<S>(C.q2, L.at, C.Z); gap_close_arr<S>(C.w, L.at, C.Z); gap_close_arr<S>(C.wp, L.at, C.Z);
      gap_close_arr<S>(C.wm, L.at, C.Z); gap_close_arr<S>(C.rr, L.at, C.Z); gap_close_arr<S>(C.n, L.at, C.Z);
    }
  }
  C.Z = L.znew;
  if (L.p0 >= 0 || L.w0i >= 0) {
    const float q0 = q2_of(L.p0x, L.p0y, L.p0z), q1 = q2_of(L.p1x, L.p1y, L.p1z);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      const int i = s * 64 + lane;
      if (i == L.p0) { C.n[s] = L.p0N; C.qx[s] = L.p0x; C.qy[s] = L.p0y; C.qz[s] = L.p0z; C.q2[s] = q0; }
      if (i == L.p1) { C.n[s] = L.p1N; C.qx[s] = L.p1x; C.qy[s] = L.p1y; C.qz[s] = L.p1z; C.q2[s] = q1; }
      if (i == L.w0i) C.w[s] = L.w0;
      if (i == L.w1i) C.w[s] = L.w1;
    }
  }
  refresh<S>(C, T, cd, L.lo, L.hi);
}

struct Header {
  int Z, nk, ens, stall, local;
  uint32_t cid;
  int64_t base, next_sample;
  double t;
  uint64_t ev, zsum;
  float dt_prev;
};

__device__ __forceinline__ void load_header(Header& H, const ChainArrays& ch, int c) {
  H.Z = ch.Z[c]; H.nk = ch.nk[c]; H.ens = ch.ens[c]; H.stall = ch.stall[c];
  H.local = ch.local[c]; H.cid = ch.cid[c]; H.base = ch.base[c]; H.next_sample = ch.next_sample[c];
  H.t = ch.t[c]; H.ev = ch.ev[c]; H.zsum = ch.zsum[c]; H.dt_prev = ch.dt_prev[c];
}
__device__ __forceinline__ void store_header(const Header& H, const ChainArrays& ch, int c) {
  if (lane_id() == 0) {
    ch.Z[c] = H.Z; ch.stall[c] = H.stall; ch.next_sample[c] = H.next_sample; ch.t[c] = H.t;
    ch.ev[c] = H.ev; ch.zsum[c] = H.zsum; ch.dt_prev[c] = H.dt_prev;
  }
}

template <int S>
__device__ __forceinline__ void load_strands(Chain<S>& C, const ChainArrays& ch, int64_t base) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int i = s * 64 + lane;
    if (i < C.Z) {
      const int64_t o = base + i;
      C.qx[s] = ch.Qx[o]; C.qy[s] = ch.Qy[o]; C.qz[s] = ch.Qz[o]; C.n[s] = ch.N[o]; C.w[s] = ch.w[o];
    } else {
      C.qx[s] = 0.0f; C.qy[s] = 0.0f; C.qz[s] = 0.0f; C.n[s] = 0; C.w[s] = 0.0f;
    }
    C.q2[s] = q2_of(C.qx[s], C.qy[s], C.qz[s]);
    C.wp[s] = 0.0f; C.wm[s] = 0.0f; C.rr[s] = 0.0f;
  }
}

template <int S>
__device__ __forceinline__ void store_strands(const Chain<S>& C, const ChainArrays& ch, int64_t base, int lo,
                                              int hi) {
  const int lane = lane_id();
#pragma unroll
  for (int s = 0; s < S; ++s) {
    const int i = s * 64 + lane;
    if (i >= lo && i <= hi && i < C.Z) {
      const int64_t o = base + i;
      ch.Qx[o] = C.qx[s]; ch.Qy[o] = C.qy[s]; ch.Qz[o] = C.qz[s]; ch.N[o] = C.n[s]; ch.w[o] = C.w[s];
    }
  }
}

// Table base for a chain of ensemble `ens`: the wave's LDS copy when the launch
// stages tables in LDS (one copy per wave, refreshed when the ensemble changes),
// otherwise the global arena.
__device__ __forceinline__ Tabs wave_tables(const KmcArgs& a, const EnsParams& P, int ens, float* lds_wave,
                                            int& lds_ens) {
  Tabs T;
  T.n = P.nmax + 2;
  if (lds_wave) {
    if (lds_ens != ens) {
      const int tsz = 9 * T.n + 4 * P.nmax + 4;
      const float* src = a.tabs + P.tab_off;
      for (int i = lane_id(); i < tsz; i += 64) lds_wave[i] = src[i];
      __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
      __builtin_amdgcn_wave_barrier();
      __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
      lds_ens = ens;
    }
    T.b = lds_wave;
  } else {
    T.b = a.tabs + P.tab_off;
  }
  return T;
}

// Runs chain c until its clock passes the ensemble's t_end (or its jump counter
// reaches ev_stop, or the chain outgrows 64*S strands).  Returns 1 on overflow.
template <int S, bool PLAIN>
__device__ __forceinline__ int run_chain(const KmcArgs& a, int c, float* lds_wave, int& lds_ens) {
  Header H;
  load_header(H, a.ch, c);
  if (H.stall) return 0;
  if (H.Z > 64 * S) return 1;
  EnsLocal E;
  load_ens(E, a.ens, H.ens);
  const uint64_t ev_stop = a.use_ev_stop ? a.ch.ev_stop[c] : ~0ull;
  // nothing to do: clock at target, no sample pending, no deformation pending
  {
    const bool pending_eq = E.sample_mode == kSampleEq && H.next_sample < E.k_end &&
                            H.t > (double)H.next_sample * E.delta;
    const bool pending_fl = E.sample_mode == kSampleFlow && H.next_sample < E.k_end &&
                            H.t >= (double)H.next_sample * E.delta;
    const bool pending_deform = E.flow && H.dt_prev != 0.0f;
    if ((H.t >= E.t_end || H.ev >= ev_stop) && !pending_eq && !pending_fl && !pending_deform) return 0;
  }
  const EnsParams& P = a.ens[H.ens];
  const Tabs T = wave_tables(a, P, H.ens, lds_wave, lds_ens);
  const float e1 = 1.0f / (P.beta * (float)H.nk);
  Chain<S> C;
  C.Z = H.Z;
  load_strands<S>(C, a.ch, H.base);
  refresh<S>(C, T, E.cd, 0, C.Z - 2);
  RandCache RC;
  RC.valid = false;
  RC.base = 0;
  int dlo = 1 << 30, dhi = -1;
  int overflow = 0;
  int done = 0;
  for (;;) {
    if (E.flow) {
      if (H.dt_prev != 0.0f) {
        deform<S>(C, E.P, H.dt_prev);
        refresh<S>(C, T, E.cd, 0, C.Z - 2);
        dlo = 0; dhi = C.Z - 1;
      }
      H.dt_prev = 0.0f;
    }
    record_samples<S>(C, T, E, H.t, H.next_sample, H.local, a.stage);
    if (H.stall || H.t >= E.t_end || H.ev >= ev_stop) break;
    if (PLAIN && done >= 1) break;
    if (C.Z >= 64 * S) { overflow = 1; break; }
    const int zb = C.Z;
    Plan L;
    kmc_pick<S>(C, T, E, e1, H.cid, RC, H.t, H.ev, H.dt_prev, H.stall, L);
    if (L.kind == kEvStall) break;
    H.zsum += (uint64_t)zb;
    ++done;
    if (L.kind != kEvNone) {
      kmc_apply<S>(C, T, E.cd, L);
      dlo = min(dlo, L.dlo);
      dhi = max(dhi, L.dhi);
    }
  }
  H.Z = C.Z;
  if (PLAIN) store_strands<S>(C, a.ch, H.base, dlo, dhi);
  else store_strands<S>(C, a.ch, H.base, 0, C.Z - 1);
  store_header(H, a.ch, c);
  return overflow;
}

template <int S>
__global__ void __launch_bounds__(256) kmc_batched_kernel(KmcArgs a) {
  extern __shared__ float lds_tabs[];
  float* lds_wave = a.lds_floats > 0 ? lds_tabs + (threadIdx.x >> 6) * a.lds_floats : nullptr;
  int lds_ens = -1;
  for (;;) {
    int item = 0;
    if (lane_id() == 0) item = atomicAdd(a.queue, 1);
    item = __builtin_amdgcn_readfirstlane(item);
    if (item >= *a.n_work) break;
    const int c = a.work[item];
    if (run_chain<S, false>(a, c, lds_wave, lds_ens)) {
      if (lane_id() == 0) a.overflow[atomicAdd(a.n_overflow, 1)] = c;
    }
  }
}

template <int S>
__global__ void __launch_bounds__(256) kmc_plain_kernel(KmcArgs a) {
  const int item = __builtin_amdgcn_readfirstlane((int)((blockIdx.x * blockDim.x + threadIdx.x) >> 6));
  if (item >= *a.n_work) return;
  int lds_ens = -1;
  const int c = a.work[item];
  if (run_chain<S, true>(a, c, nullptr, lds_ens)) {
    if (lane_id() == 0) a.overflow[atomicAdd(a.n_overflow, 1)] = c;
  }
}

}  // namespace dsm
