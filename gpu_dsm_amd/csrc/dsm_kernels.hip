// dsm_kernels.hip -- MI355X (gfx950) kernels for the Discrete Slip-Link Model ensemble KMC path.
//
// One wave64 owns one chain.  Strand i sits in lane i%64 of register slot i/64
// (S slots per lane, S a template parameter, so a chain of up to 64*S strands
// lives in VGPRs).  A KMC jump (ktaletsk/gpu_dsm common/eq_ensemble_kernel.cu:242-663,
// common/ensemble_kernel.cu:188-502) is:
//   1. per-pair rates (sliding + constraint dynamics) cached per lane; the five
//      chain-end rates summed into the last position,
//   2. a wave64 inclusive scan (DPP tree) -> total W, dt = 1/W, FP64 clock += dt,
//   3. pr = W*u from Philox(seed, chain, event); the pair is the first lane
//      whose inclusive sum exceeds pr (ballot + ffs); the sub-event follows the
//      reference order CDd, SD+, SD-, CDc (pairs) / CDc tail, SDc head/tail,
//      SDd head/tail (ends),
//   4. the effect: N updates for shifts; split/merge as in-wave shifts of all
//      strand records; only the pair rates next to the change are recomputed.
// Two forms share this code:
//   * plain:   one jump per chain per launch, state read from and written to HBM
//              (the reference's one-step-per-launch structure);
//   * batched: persistent waves pull chains from a work queue, keep the chain in
//              registers for many jumps (until its clock passes the window end),
//              then write it back once.
#include "dsm_kmc.h"
#include "dsm_launch.h"

namespace dsm {

// ------------------------------------------------------------------ equilibrium init
// Thread per chain.  Inner Kuhn-step boundary i (1 <= i < nk) carries an
// entanglement iff Philox word (i-1) < 2^32/(1+beta): Z-1 ~ Binomial(nk-1, 1/(1+beta))
// and N|Z uniform over compositions (the laws of common/chain.cu:38-88).
// Internal strands get Q ~ N(0, N/3) per component (common/chain.cu:90-102),
// links get CD rates from f_t (common/chain.cu:155-158), with the partner length of
// each link drawn from the background distribution when there is one (:141-149).
__global__ void __launch_bounds__(256) init_eq_kernel(ChainArrays ch, const EnsParams* ens, const float* tabs,
                                                        const int* chains, int n) {
  const int idx = blockIdx.x * blockDim.x + threadIdx.x;
  if (idx >= n) return;
  const int c = chains[idx];
  const EnsParams& P = ens[ch.ens[c]];
  const uint32_t k0 = P.key0, k1 = P.key1;
  const int nk = ch.nk[c];
  const uint32_t cid = ch.cid[c];
  const int64_t b = ch.base[c];
  int Z = 0, len = 0;
  uint4 x = make_uint4(0, 0, 0, 0);
  for (int i = 1; i <= nk; ++i) {
    ++len;
    if (i == nk) { ch.N[b + Z] = len; ++Z; break; }
    const int wi = i - 1;
    if ((wi & 3) == 0) x = philox4x32_10(make_uint4((uint32_t)(wi >> 2), 0u, cid, kPurposeInitCut), k0, k1);
    const uint32_t word = (wi & 3) == 0 ? x.x : (wi & 3) == 1 ? x.y : (wi & 3) == 2 ? x.z : x.w;
    if (word < P.cut_threshold) { ch.N[b + Z] = len; ++Z; len = 0; }
  }
  for (int k = 0; k < Z; ++k) {
    float qx = 0.0f, qy = 0.0f, qz = 0.0f;
    if (k > 0 && k < Z - 1) {
      const uint4 g = philox4x32_10(make_uint4((uint32_t)k, 0u, cid, kPurposeInitQ), k0, k1);
      float g0, g1, g2, g3;
      boxmuller(g.x, g.y, g0, g1);
      boxmuller(g.z, g.w, g2, g3);
      const float sig = sqrtf((float)ch.N[b + k] / 3.0f);
      qx = g0 * sig; qy = g1 * sig; qz = g2 * sig;
    }
    ch.Qx[b + k] = qx; ch.Qy[b + k] = qy; ch.Qz[b + k] = qz;
    ch.w[b + k] = 0.0f;
  }
  if (P.cd) {
    const float* pd = tabs + P.pd_off;
    uint4 y = make_uint4(0, 0, 0, 0);
    for (int k = 0; k < Z - 1; ++k) {
      if ((k & 3) == 0) x = philox4x32_10(make_uint4((uint32_t)(k >> 2), 0u, cid, kPurposeInitTau), k0, k1);
      const uint32_t word = (k & 3) == 0 ? x.x : (k & 3) == 1 ? x.y : (k & 3) == 2 ? x.z : x.w;
      if (P.partner) {
        if ((k & 3) == 0) y = philox4x32_10(make_uint4((uint32_t)(k >> 2), 0u, cid, kPurposeInitPartner), k0, k1);
        const uint32_t pw = (k & 3) == 0 ? y.x : (k & 3) == 1 ? y.y : (k & 3) == 2 ? y.z : y.w;
        ch.w[b + k] = rate_ft_partner(pd, (int)(pw >> (32 - kPartnerBits)), u01(word));
      } else {
        ch.w[b + k] = rate_ft(&P, u01(word));
      }
    }
  }
  ch.Z[c] = Z; ch.stall[c] = 0; ch.t[c] = 0.0; ch.ev[c] = 0; ch.dt_prev[c] = 0.0f; ch.zsum[c] = 0;
  ch.next_sample[c] = (P.sample_mode == kSampleFlow) ? 1 : 0;
}

// ------------------------------------------------------------------ utility kernels
__global__ void fill_u32_kernel(uint32_t* p, uint32_t v, int64_t n) {
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = v;
}

// one block: out[0] += sum of ev, out[1] += sum of zsum over a chain range
__global__ void __launch_bounds__(1024) sum_counters_kernel(const uint64_t* ev, const uint64_t* zsum, int chain0,
                                                             int n, unsigned long long* out) {
  __shared__ unsigned long long r0[1024], r1[1024];
  unsigned long long a0 = 0, a1 = 0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    a0 += ev[chain0 + i];
    a1 += zsum[chain0 + i];
  }
  r0[threadIdx.x] = a0;
  r1[threadIdx.x] = a1;
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) { r0[threadIdx.x] += r0[threadIdx.x + h]; r1[threadIdx.x] += r1[threadIdx.x + h]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { atomicAdd(out, r0[0]); atomicAdd(out + 1, r1[0]); }
}

__global__ void set_ev_stop_kernel(const uint64_t* ev, uint64_t* ev_stop, int chain0, int n, uint64_t add) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) ev_stop[chain0 + i] = ev[chain0 + i] + add;
}

// Chains of a pass into per-geometry lists by their current Z.  Geometry b takes a chain
// whose Z is below its capacity -- for the wave-form geometries (b >= 2) below 7/8 of it,
// so that a long chain, whose Z drifts by a few tens over a window, rarely outgrows its
// geometry mid-pass (it would then finish alone in the late round, after everything
// else; measured on the polydisperse melt, the same margin at the row and pair
// boundaries costs more in the first round than it saves in the late one).  Stalled
// chains go to geometry 0, which skips them.  Lists are cap entries apart, counts[b]
// their lengths.
// One atomic per wave and geometry: the lanes of a geometry get consecutive slots.
// margin_div > 0: a chain within capacity/margin_div of a wave-form geometry's capacity
// starts in the next one (it would likely outgrow it during the pass)
__global__ void split_geoms_kernel(const int* src, int n, const int* Z, const int* stall, int* lists, int cap,
                                   int* counts, int margin_div) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  const int c = i < n ? src[i] : -1;
  int b = -1;
  if (c >= 0) {
    const int z = Z[c];
    b = 0;
    if (stall[c] == 0)
      while (b + 1 < kNumGeoms &&
             z >= kmc_geom_capacity(b) - (b >= 2 && margin_div > 0 ? kmc_geom_capacity(b) / margin_div : 0))
        ++b;
  }
  const int lane = (int)__lane_id();
  const uint64_t below = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
#pragma unroll
  for (int g = 0; g < kNumGeoms; ++g) {
    const uint64_t m = __ballot(b == g);
    if (m == 0) continue;
    const int leader = (int)__builtin_ctzll(m);
    int base = 0;
    if (lane == leader) base = atomicAdd(counts + g, __popcll(m));
    base = __shfl(base, leader);
    if (b == g) lists[(int64_t)g * cap + base + __popcll(m & below)] = c;
  }
}

// Row-form work list in decreasing Z (16 buckets of four; Z >= 64 first, stalled chains
// last): the chains started last, at the end of a pass, are the shortest ones.
__device__ __forceinline__ int zbucket(const int* Z, const int* stall, int c) {
  if (stall[c] != 0) return 16;
  const int z = Z[c];
  return z >= 64 ? 0 : 15 - (z >> 2);
}
constexpr int kZBuckets = 17;

// Block-local histograms in LDS, one global atomic per block and bucket (a per-element
// global atomic on 17 counters serialises in L2).  Grid-stride over the elements.
// src == null: the chains first .. first+n-1 (one ensemble's range)
__global__ void __launch_bounds__(1024) zhist_kernel(const int* src, int first, int n, const int* Z, const int* stall,
                                                     int* hist) {
  __shared__ int lh[kZBuckets];
  if (threadIdx.x < kZBuckets) lh[threadIdx.x] = 0;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = src ? src[i] : first + i;
    if (c >= 0) atomicAdd(lh + zbucket(Z, stall, c), 1);
  }
  __syncthreads();
  if (threadIdx.x < kZBuckets && lh[threadIdx.x]) atomicAdd(hist + threadIdx.x, lh[threadIdx.x]);
}

// hist -> exclusive offsets (in place), total into *n_out
__global__ void zscan_kernel(int* hist, int* n_out) {
  if (threadIdx.x != 0) return;
  int acc = 0;
  for (int g = 0; g < kZBuckets; ++g) { const int h = hist[g]; hist[g] = acc; acc += h; }
  *n_out = acc;
}

// Each block reserves its range of every bucket with one atomic, then places its
// elements with LDS atomics (order within a bucket is arbitrary).
__global__ void __launch_bounds__(1024) zscatter_kernel(const int* src, int first, int n, const int* Z,
                                                        const int* stall, int* cursor, int* dst) {
  __shared__ int lh[kZBuckets], base[kZBuckets];
  if (threadIdx.x < kZBuckets) lh[threadIdx.x] = 0;
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = src ? src[i] : first + i;
    if (c >= 0) atomicAdd(lh + zbucket(Z, stall, c), 1);
  }
  __syncthreads();
  if (threadIdx.x < kZBuckets) {
    base[threadIdx.x] = lh[threadIdx.x] ? atomicAdd(cursor + threadIdx.x, lh[threadIdx.x]) : 0;
    lh[threadIdx.x] = 0;
  }
  __syncthreads();
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = src ? src[i] : first + i;
    if (c >= 0) {
      const int b = zbucket(Z, stall, c);
      dst[base[b] + atomicAdd(lh + b, 1)] = c;
    }
  }
}

// counts of a sequential pass: cnt[0] = n0, the rest (lists, queues, order scratch) zero
__global__ void init_counts_kernel(int* cnt, int n, int n0) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) cnt[i] = i == 0 ? n0 : 0;
}

__global__ void iota_kernel(int* p, int start, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < n) p[i] = start + i;
}

// flow stress: per (sample, ensemble) FP64 sums of the 8 components over healthy
// chains (stall == 0, finite, recorded); out[...*9 + 8] = chain count.
// (common/ensemble_call_block.cu:452-505 averages in FP32 on the host.)
// A row of a chain counts when the chain recorded it (Z column >= 1: staging rows start
// at zero and a chain that has stalled records nothing after the stall) and it holds no
// NaN, so a chain that stalls mid-window keeps the rows it recorded before, as the
// reference excludes a stalled chain only from the outputs after its stall
// (common/ensemble_call_block.cu:471-491, per output step).
__global__ void __launch_bounds__(256) flow_reduce_kernel(const float* stage, int nchains, int nrows, double* out) {
  const int row = blockIdx.x;
  if (row >= nrows) return;
  __shared__ double red[9][256];
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = 0.0;
  for (int c = threadIdx.x; c < nchains; c += blockDim.x) {
    const float* o = stage + ((int64_t)row * nchains + c) * 8;
    bool ok = o[7] > 0.0f;
#pragma unroll
    for (int q = 0; q < 8; ++q) ok = ok && (o[q] == o[q]);
    if (ok) {
#pragma unroll
      for (int q = 0; q < 8; ++q) acc[q] += (double)o[q];
      acc[8] += 1.0;
    }
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
#pragma unroll
      for (int q = 0; q < 9; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 9; ++q) out[(int64_t)row * 9 + q] = red[q][0];
  }
}

// ------------------------------------------------------------------ multi-tau correlator
// Thread per chain (lanes = consecutive chains -> every access below is coalesced).
// Per-chain FP64 state: D[level][slot][3], A[level][3], C[level][lag]
// (C accumulates the sum of the three component products).  The level
// counters are identical for all chains (sampling is synchronous) and are
// passed in; the host advances its copy by the same rule after each launch.
// Insert rule, block averages and lag products: common/eq_ensemble_kernel.cu:133-227.
struct CorrArgs {
  const float* stage;   // [k][chain][3]
  double* D;            // [((lvl*p + slot)*3 + comp)*nch + chain]
  double* A;            // [(lvl*3 + comp)*nch + chain]
  double* C;            // [(lvl*p + lag)*nch + chain]
  const int* stall;     // global stall flags
  int chain0, nch, p, m, s, nrows;
  const int* cnt;       // [3*s]: ins, macc, nins (clamped to p) at window start
  int type;             // 0: stress (block averages, products); 1: MSD (first value, squared differences)
};

// One value into level `lvl` and, through the block pushes, the levels above it
// (D, A and C in device memory).
__device__ __forceinline__ void corr_insert_from(const CorrArgs& a, int c, int lvl, double vx, double vy, double vz,
                                                 int* ins, int* macc, int* nins) {
  const int64_t nch = a.nch;
  const int p = a.p, m = a.m;
  for (;;) {
    if (lvl == a.s) break;
    const int in = ins[lvl];
    double* D = a.D + (int64_t)(lvl * p) * 3 * nch + c;
    D[(int64_t)(in * 3 + 0) * nch] = vx;
    D[(int64_t)(in * 3 + 1) * nch] = vy;
    D[(int64_t)(in * 3 + 2) * nch] = vz;
    double* A = a.A + (int64_t)(lvl * 3) * nch + c;
    // MSD carries the first value of each block up unaveraged
    // (common/eq_ensemble_kernel.cu:181-189)
    if (a.type == 0 || macc[lvl] == 0) { A[0] += vx; A[nch] += vy; A[2 * nch] += vz; }
    macc[lvl]++;
    const bool push = macc[lvl] == m;
    double ux = 0.0, uy = 0.0, uz = 0.0;
    if (push) {
      if (a.type == 0) { ux = A[0] / m; uy = A[nch] / m; uz = A[2 * nch] / m; }
      else { ux = A[0]; uy = A[nch]; uz = A[2 * nch]; }
      A[0] = 0.0; A[nch] = 0.0; A[2 * nch] = 0.0;
      macc[lvl] = 0;
    }
    const int jmin = (lvl == 0) ? 0 : p / m;
    double* Cc = a.C + (int64_t)(lvl * p) * nch + c;
    for (int j = jmin; j < p; ++j) {
      if (j > nins[lvl]) break;
      int idx = in - j;
      if (idx < 0) idx += p;
      const double dx = D[(int64_t)(idx * 3 + 0) * nch];
      const double dy = D[(int64_t)(idx * 3 + 1) * nch];
      const double dz = D[(int64_t)(idx * 3 + 2) * nch];
      if (a.type == 0) {
        Cc[(int64_t)j * nch] += vx * dx + vy * dy + vz * dz;
      } else {
        const double ex = vx - dx, ey = vy - dy, ez = vz - dz;
        Cc[(int64_t)j * nch] += ex * ex + ey * ey + ez * ez;
      }
    }
    ins[lvl] = (in + 1 == p) ? 0 : in + 1;
    nins[lvl] = nins[lvl] < p ? nins[lvl] + 1 : p;
    if (!push) break;
    ++lvl;
    vx = ux; vy = uy; vz = uz;
  }
}

// As corr_insert_from for p = P, with the lag loop in groups of four lags whose shift
// register values and lag sums are loaded together: the loads of a group are in flight
// at once instead of one lag after another (the lag sums and the shift register are
// distinct arrays, which the compiler cannot assume).  Same operations in the same order.
template <int P>
__device__ __forceinline__ void corr_insert_from_p(const CorrArgs& a, int c, int lvl, double vx, double vy, double vz,
                                                   int* ins, int* macc, int* nins) {
  const int64_t nch = a.nch;
  const int m = a.m;
  for (;;) {
    if (lvl == a.s) break;
    const int in = ins[lvl];
    const int nin = nins[lvl];
    double* __restrict__ D = a.D + (int64_t)(lvl * P) * 3 * nch + c;
    double* __restrict__ A = a.A + (int64_t)(lvl * 3) * nch + c;
    double* __restrict__ Cc = a.C + (int64_t)(lvl * P) * nch + c;
    const int jmin = (lvl == 0) ? 0 : P / m;
    // lag j pairs v with the value inserted j steps before it (j = 0: v itself)
#pragma unroll
    for (int j0 = 0; j0 < P; j0 += 4) {
      if (j0 + 3 < jmin || j0 > nin) continue;
      double dx[4], dy[4], dz[4], cv[4];
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int j = j0 + k;
        int idx = in - j;
        if (idx < 0) idx += P;
        const bool use = j >= jmin && j <= nin;
        dx[k] = vx; dy[k] = vy; dz[k] = vz; cv[k] = 0.0;
        if (use && j > 0) {
          dx[k] = D[(int64_t)(idx * 3 + 0) * nch];
          dy[k] = D[(int64_t)(idx * 3 + 1) * nch];
          dz[k] = D[(int64_t)(idx * 3 + 2) * nch];
        }
        if (use) cv[k] = Cc[(int64_t)j * nch];
      }
#pragma unroll
      for (int k = 0; k < 4; ++k) {
        const int j = j0 + k;
        if (!(j >= jmin && j <= nin)) continue;
        if (a.type == 0) {
          cv[k] += vx * dx[k] + vy * dy[k] + vz * dz[k];
        } else {
          const double ex = vx - dx[k], ey = vy - dy[k], ez = vz - dz[k];
          cv[k] += ex * ex + ey * ey + ez * ez;
        }
        Cc[(int64_t)j * nch] = cv[k];
      }
    }
    D[(int64_t)(in * 3 + 0) * nch] = vx;
    D[(int64_t)(in * 3 + 1) * nch] = vy;
    D[(int64_t)(in * 3 + 2) * nch] = vz;
    // MSD carries the first value of each block up unaveraged
    // (common/eq_ensemble_kernel.cu:181-189)
    double ax = A[0], ay = A[nch], az = A[2 * nch];
    if (a.type == 0 || macc[lvl] == 0) { ax += vx; ay += vy; az += vz; }
    macc[lvl]++;
    const bool push = macc[lvl] == m;
    double ux = 0.0, uy = 0.0, uz = 0.0;
    if (push) {
      if (a.type == 0) { ux = ax / m; uy = ay / m; uz = az / m; }
      else { ux = ax; uy = ay; uz = az; }
      ax = 0.0; ay = 0.0; az = 0.0;
      macc[lvl] = 0;
    }
    A[0] = ax; A[nch] = ay; A[2 * nch] = az;
    ins[lvl] = (in + 1 == P) ? 0 : in + 1;
    nins[lvl] = nin < P ? nin + 1 : P;
    if (!push) break;
    ++lvl;
    vx = ux; vy = uy; vz = uz;
  }
}

// Any p: every level in device memory.
__global__ void __launch_bounds__(256) corr_update_kernel(CorrArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nch) return;
  int ins[40], macc[40], nins[40];
  for (int k = 0; k < a.s; ++k) { ins[k] = a.cnt[k]; macc[k] = a.cnt[a.s + k]; nins[k] = a.cnt[2 * a.s + k]; }
  const int64_t nch = a.nch;
  for (int row = 0; row < a.nrows; ++row) {
    const float* o = a.stage + ((int64_t)row * nch + c) * 3;
    corr_insert_from(a, c, 0, o[0], o[1], o[2], ins, macc, nins);
  }
}

// p = P (16 in every reference input): level 0, which takes every sample row, is held
// in registers for the whole window -- its P slots in age order (h[0] the newest), its P
// lag sums and its block accumulator -- and written back once; the levels above take
// the block pushes in device memory as before.  Every sum is formed in the same order
// as in corr_update_kernel, so both give identical results.
template <int P>
__global__ void __launch_bounds__(256) __attribute__((amdgpu_waves_per_eu(2))) corr_update_reg_kernel(CorrArgs a) {
  const int c = blockIdx.x * blockDim.x + threadIdx.x;
  if (c >= a.nch) return;
  int ins[40], macc[40], nins[40];
  for (int k = 1; k < a.s; ++k) { ins[k] = a.cnt[k]; macc[k] = a.cnt[a.s + k]; nins[k] = a.cnt[2 * a.s + k]; }
  int in0 = a.cnt[0], macc0 = a.cnt[a.s], nins0 = a.cnt[2 * a.s];
  const int64_t nch = a.nch;
  const int m = a.m;
  double* D = a.D + c;
  double* A = a.A + c;
  double* Cc = a.C + c;
  double hx[P], hy[P], hz[P], cs[P];
#pragma unroll
  for (int k = 0; k < P; ++k) {
    const int slot = (in0 - 1 - k + 2 * P) % P;
    hx[k] = D[(int64_t)(slot * 3 + 0) * nch];
    hy[k] = D[(int64_t)(slot * 3 + 1) * nch];
    hz[k] = D[(int64_t)(slot * 3 + 2) * nch];
    cs[k] = Cc[(int64_t)k * nch];
  }
  double ax = A[0], ay = A[nch], az = A[2 * nch];
  for (int row = 0; row < a.nrows; ++row) {
    const float* o = a.stage + ((int64_t)row * nch + c) * 3;
    const double vx = o[0], vy = o[1], vz = o[2];
    if (a.type == 0 || macc0 == 0) { ax += vx; ay += vy; az += vz; }
    ++macc0;
    const bool push = macc0 == m;
    double ux = 0.0, uy = 0.0, uz = 0.0;
    if (push) {
      if (a.type == 0) { ux = ax / m; uy = ay / m; uz = az / m; }
      else { ux = ax; uy = ay; uz = az; }
      ax = 0.0; ay = 0.0; az = 0.0;
      macc0 = 0;
    }
    // lag j pairs the new value with the one inserted j steps before it (h[j-1])
#pragma unroll
    for (int j = 0; j < P; ++j) {
      const double dx = j == 0 ? vx : hx[j - 1 < 0 ? 0 : j - 1];
      const double dy = j == 0 ? vy : hy[j - 1 < 0 ? 0 : j - 1];
      const double dz = j == 0 ? vz : hz[j - 1 < 0 ? 0 : j - 1];
      double add;
      if (a.type == 0) {
        add = vx * dx + vy * dy + vz * dz;
      } else {
        const double ex = vx - dx, ey = vy - dy, ez = vz - dz;
        add = ex * ex + ey * ey + ez * ez;
      }
      if (j <= nins0) cs[j] += add;
    }
#pragma unroll
    for (int k = P - 1; k > 0; --k) { hx[k] = hx[k - 1]; hy[k] = hy[k - 1]; hz[k] = hz[k - 1]; }
    hx[0] = vx; hy[0] = vy; hz[0] = vz;
    in0 = (in0 + 1 == P) ? 0 : in0 + 1;
    nins0 = nins0 < P ? nins0 + 1 : P;
    if (push) corr_insert_from_p<P>(a, c, 1, ux, uy, uz, ins, macc, nins);
  }
  // write back: the slots of this window's values (the newest min(nrows, P)), the lag
  // sums and the block accumulator
#pragma unroll
  for (int k = 0; k < P; ++k) {
    if (k < a.nrows) {
      const int slot = (in0 - 1 - k + 2 * P) % P;
      D[(int64_t)(slot * 3 + 0) * nch] = hx[k];
      D[(int64_t)(slot * 3 + 1) * nch] = hy[k];
      D[(int64_t)(slot * 3 + 2) * nch] = hz[k];
    }
    Cc[(int64_t)k * nch] = cs[k];
  }
  A[0] = ax; A[nch] = ay; A[2 * nch] = az;
}

// G(level, lag) = sum over healthy chains of C / (3 * n_pairs * n_healthy); the chains
// are first .. first+count-1 of an ensemble of ld chains (C is [point][ld])
__global__ void __launch_bounds__(256) corr_reduce_kernel(const double* C, int ld, const int* stall, int chain0,
                                                            int first, int count, int npts, double* out_sum,
                                                            int* out_healthy) {
  const int pt = blockIdx.x;
  if (pt >= npts) return;
  __shared__ double red[256];
  __shared__ int cnt[256];
  double acc = 0.0;
  int h = 0;
  for (int c = first + threadIdx.x; c < first + count; c += blockDim.x) {
    if (stall[chain0 + c] == 0) { acc += C[(int64_t)pt * ld + c]; ++h; }
  }
  red[threadIdx.x] = acc;
  cnt[threadIdx.x] = h;
  __syncthreads();
  for (int s = blockDim.x / 2; s > 0; s >>= 1) {
    if ((int)threadIdx.x < s) { red[threadIdx.x] += red[threadIdx.x + s]; cnt[threadIdx.x] += cnt[threadIdx.x + s]; }
    __syncthreads();
  }
  if (threadIdx.x == 0) { out_sum[pt] = red[0]; if (pt == 0) *out_healthy = cnt[0]; }
}

// ------------------------------------------------------------------ captured passes
// Head of a captured window pass: the window's parameters, which change every window,
// come from a block of pinned host memory the host fills before each replay, so the
// graph itself never changes: (t_end, k0, k_end) of every ensemble of the pass into its
// EnsParams, then the correlator level counters.  One block.
__global__ void __launch_bounds__(256) stage_window_kernel(const WinParam* __restrict__ wp, int n_ens,
                                                           EnsParams* P, const int* __restrict__ cnt_src,
                                                           int* cnt_dst, int n_cnt) {
  for (int i = threadIdx.x; i < n_ens; i += blockDim.x) {
    const WinParam w = wp[i];
    P[w.ens].t_end = w.t_end;
    P[w.ens].k0 = w.k0;
    P[w.ens].k_end = w.k_end;
  }
  for (int i = threadIdx.x; i < n_cnt; i += blockDim.x) cnt_dst[i] = cnt_src[i];
}

// ------------------------------------------------------------------ ensemble stress now
// Stress of the current state (common/ensemble_kernel.cu:504-538 per chain, averaged by
// common/ensemble.cu:210-221): per chain sigma_ab = -sum_i 3 Q_ia Q_ib / N_i, Lpp =
// sum_i |Q_i| and Z, in FP64 from the stored FP32 strands, after the deformation still
// pending from the chain's last jump (dt_prev; the reference applies it with olddt).
// Chains that have stalled are left out.  A fixed grid of kStressBlocks blocks strides
// over the chains; block b writes its partial sums to part[9b .. 9b+8] (xx yy zz xy yz zx
// Lpp Z count), which the host adds in block order, so the result does not depend on
// timing.
__global__ void __launch_bounds__(256) chain_stress_kernel(ChainArrays ch, int chain0, int n, KappaArg K,
                                                           double* part) {
  __shared__ double red[9][256];
  double acc[9];
#pragma unroll
  for (int q = 0; q < 9; ++q) acc[q] = 0.0;
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = chain0 + i;
    if (ch.stall[c] != 0) continue;
    const int Z = ch.Z[c];
    const int64_t b = ch.base[c];
    const double dt = (double)ch.dt_prev[c];
    double s[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = 0; k < Z; ++k) {
      double x = ch.Qx[b + k], y = ch.Qy[b + k], z = ch.Qz[b + k];
      if (dt != 0.0) {
        const double x1 = x + dt * ((double)K.k[0] * x + (double)K.k[1] * y + (double)K.k[2] * z);
        const double y1 = y + dt * ((double)K.k[3] * x + (double)K.k[4] * y + (double)K.k[5] * z);
        const double z1 = z + dt * ((double)K.k[6] * x + (double)K.k[7] * y + (double)K.k[8] * z);
        x = x1; y = y1; z = z1;
      }
      const double t3 = 3.0 / (double)ch.N[b + k];
      s[0] -= t3 * x * x; s[1] -= t3 * y * y; s[2] -= t3 * z * z;
      s[3] -= t3 * x * y; s[4] -= t3 * y * z; s[5] -= t3 * x * z;
      s[6] += sqrt(x * x + y * y + z * z);
    }
#pragma unroll
    for (int q = 0; q < 7; ++q) acc[q] += s[q];
    acc[7] += (double)Z;
    acc[8] += 1.0;
  }
#pragma unroll
  for (int q = 0; q < 9; ++q) red[q][threadIdx.x] = acc[q];
  __syncthreads();
  for (int h = blockDim.x / 2; h > 0; h >>= 1) {
    if ((int)threadIdx.x < h) {
#pragma unroll
      for (int q = 0; q < 9; ++q) red[q][threadIdx.x] += red[q][threadIdx.x + h];
    }
    __syncthreads();
  }
  if (threadIdx.x == 0) {
#pragma unroll
    for (int q = 0; q < 9; ++q) part[blockIdx.x * 9 + q] = red[q][0];
  }
}

// ------------------------------------------------------------------ state invariants
// Counts, over chains chain0 .. chain0+n-1, the violations of the model's invariants
// (App. A.1 of the chain state): out[0] Z outside 1..Nk, out[1] a strand with N < 1,
// out[2] sum of N != Nk, out[3] an end strand (index 0 or Z-1, Z > 1) with Q != 0,
// out[4] a non-finite Q or CD rate, out[5] stalled chains, out[6] chains whose clock is
// below t_min.  The counters (7 x uint64) must be zero on entry.
__global__ void __launch_bounds__(256) check_state_kernel(ChainArrays ch, int chain0, int n, double t_min,
                                                          unsigned long long* out) {
  unsigned long long v[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
    const int c = chain0 + i;
    const int Z = ch.Z[c], nk = ch.nk[c];
    if (ch.stall[c] != 0) v[5]++;
    if (!(ch.t[c] >= t_min)) v[6]++;
    if (Z < 1 || Z > nk || Z > ch.cap[c]) { v[0]++; continue; }
    const int64_t b = ch.base[c];
    int64_t sum = 0;
    bool bad_n = false, bad_f = false;
    for (int k = 0; k < Z; ++k) {
      const int N = ch.N[b + k];
      bad_n = bad_n || N < 1;
      sum += N;
      const float x = ch.Qx[b + k], y = ch.Qy[b + k], z = ch.Qz[b + k];
      bad_f = bad_f || !isfinite(x) || !isfinite(y) || !isfinite(z) || (k < Z - 1 && !isfinite(ch.w[b + k]));
    }
    if (bad_n) v[1]++;
    if (sum != nk) v[2]++;
    if (Z > 1) {
      const int64_t e0 = b, e1 = b + Z - 1;
      if (ch.Qx[e0] != 0.f || ch.Qy[e0] != 0.f || ch.Qz[e0] != 0.f || ch.Qx[e1] != 0.f || ch.Qy[e1] != 0.f ||
          ch.Qz[e1] != 0.f)
        v[3]++;
    }
    if (bad_f) v[4]++;
  }
#pragma unroll
  for (int q = 0; q < 7; ++q)
    if (v[q]) atomicAdd(out + q, v[q]);
}

// ------------------------------------------------------------------ copy probe (bandwidth)
// Plain 16-byte copy, grid-stride: the reference point for HBM byte counters and the
// achievable copy bandwidth of the device.
__global__ void copy_probe_kernel(const float4* __restrict__ in, float4* __restrict__ out, size_t n4) {
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t i = (size_t)blockIdx.x * blockDim.x + threadIdx.x; i < n4; i += stride) out[i] = in[i];
}

// ------------------------------------------------------------------ scan probe (tests)
// Canonical scan tree of the KMC kernels.  form 1: one chain of n <= 1024 positions
// (wave form, four register groups): out[0..n) inclusive sums, out[n] total.
// form 0: four chains of n <= 64 positions (row form), form 2: two chains of n <= 128
// positions (pair form), form 3: two chains of n <= 256 positions (pair form with two
// register groups), form 4: four chains of n <= 128 positions (row form with two register
// groups): chain r reads in[r*n ..] and writes out[r*(n+1) ..].
// Centre-of-mass sample of one chain state in a chosen geometry (0 row form, 1 pair form,
// 2/3/4 wave form G = 1/2/4, 7 pair form with two groups): in[4i..4i+3] = Qx Qy Qz N of
// strand i, in[4n] = Nk; out = 3 floats.
template <int CPW, int G>
__device__ void com_probe(const float* in, float* out, int n, lds_f* buf) {
  static_assert(Geo<CPW, G>::STG <= 1024, "the strand lengths fit the probe's LDS buffer");
  const LaneC<CPW> L;
  const Stage<CPW, G> S{buf, L.srow * G};
  Strands<G> C;
  int nv[4 * G];
#pragma unroll
  for (int j = 0; j < 4 * G; ++j) {
    const int p = L.pos(j);
    const bool on = p < n;
    C.qx[j] = on ? in[4 * p] : 0.0f;
    C.qy[j] = on ? in[4 * p + 1] : 0.0f;
    C.qz[j] = on ? in[4 * p + 2] : 0.0f;
    nv[j] = on ? (int)in[4 * p + 3] : 0;
    C.w[j] = 0.0f;
  }
#pragma unroll
  for (int g = 0; g < G; ++g) S.put4i(kStN, g, L.pos0, nv[4 * g], nv[4 * g + 1], nv[4 * g + 2], nv[4 * g + 3]);
  wave_lds_sync();
  float sv[3];
  com_sample<CPW, G>(C, S, n, in[4 * n], L, sv);
  if (L.lane == 0) { out[0] = sv[0]; out[1] = sv[1]; out[2] = sv[2]; }
}
__global__ void __launch_bounds__(64) com_probe_kernel(const float* in, float* out, int n, int geom) {
  __shared__ float buf[1024];  // the kStN staging array of one wave (strand lengths)
  lds_f* b = (lds_f*)buf;
  if (geom == 0) com_probe<4, 1>(in, out, n, b);
  else if (geom == 1) com_probe<2, 1>(in, out, n, b);
  else if (geom == 2) com_probe<1, 1>(in, out, n, b);
  else if (geom == 3) com_probe<1, 2>(in, out, n, b);
  else if (geom == kGeomPair2) com_probe<2, 2>(in, out, n, b);
  else if (geom == kGeomRow2) com_probe<4, 2>(in, out, n, b);
  else com_probe<1, 4>(in, out, n, b);
}

__global__ void __launch_bounds__(64) scan_probe_kernel(const float* in, float* out, int n, int form) {
  if (form == 5) {
    const LaneC<1> L;
    float v[8], g[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = L.pos(j);
      v[j] = p < n ? in[p] : 0.0f;
    }
    const float tot = chain_scan<1, 2>(v, g, n, L);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = L.pos(j);
      if (p < n) out[p] = g[j];
    }
    if (L.lane == 0) out[n] = tot;
  } else if (form == 1) {
    const LaneC<1> L;
    float v[16], g[16];
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int p = L.pos(j);
      v[j] = p < n ? in[p] : 0.0f;
    }
    const float tot = chain_scan<1, 4>(v, g, n, L);
#pragma unroll
    for (int j = 0; j < 16; ++j) {
      const int p = L.pos(j);
      if (p < n) out[p] = g[j];
    }
    if (L.lane == 0) out[n] = tot;
  } else if (form == 2) {
    const LaneC<2> L;
    const int r = L.lane >> 5;
    float v[4], g[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = L.pos(j);
      v[j] = p < n ? in[r * n + p] : 0.0f;
    }
    const float tot = chain_scan<2, 1>(v, g, n, L);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = L.pos(j);
      if (p < n) out[r * (n + 1) + p] = g[j];
    }
    if (L.sl == 0) out[r * (n + 1) + n] = tot;
  } else if (form == 4) {
    const LaneC<4> L;
    const int r = L.lane >> 4;
    float v[8], g[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = L.pos(j);
      v[j] = p < n ? in[r * n + p] : 0.0f;
    }
    const float tot = chain_scan<4, 2>(v, g, n, L);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = L.pos(j);
      if (p < n) out[r * (n + 1) + p] = g[j];
    }
    if (L.sl == 0) out[r * (n + 1) + n] = tot;
  } else if (form == 3) {
    const LaneC<2> L;
    const int r = L.lane >> 5;
    float v[8], g[8];
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = L.pos(j);
      v[j] = p < n ? in[r * n + p] : 0.0f;
    }
    const float tot = chain_scan<2, 2>(v, g, n, L);
#pragma unroll
    for (int j = 0; j < 8; ++j) {
      const int p = L.pos(j);
      if (p < n) out[r * (n + 1) + p] = g[j];
    }
    if (L.sl == 0) out[r * (n + 1) + n] = tot;
  } else {
    const LaneC<4> L;
    const int r = L.lane >> 4;
    float v[4], g[4];
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = L.pos(j);
      v[j] = p < n ? in[r * n + p] : 0.0f;
    }
    const float tot = chain_scan<4, 1>(v, g, n, L);
#pragma unroll
    for (int j = 0; j < 4; ++j) {
      const int p = L.pos(j);
      if (p < n) out[r * (n + 1) + p] = g[j];
    }
    if (L.sl == 0) out[r * (n + 1) + n] = tot;
  }
}

__global__ void math_probe_kernel(const float* in, float* out, int n, int which) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const float x = in[i];
  float r = 0.0f;
  if (which == 0) r = dexpf(x);
  else if (which == 1) r = dlogf(x);
  else if (which == 2) { float s, c; dsincos2pif(x, s, c); r = s; }
  else if (which == 3) { float s, c; dsincos2pif(x, s, c); r = c; }
  out[i] = r;
}

__global__ void philox_probe_kernel(const uint32_t* ctr_key, uint32_t* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  const uint32_t* q = ctr_key + 6 * i;
  const uint4 r = philox4x32_10(make_uint4(q[0], q[1], q[2], q[3]), q[4], q[5]);
  out[4 * i + 0] = r.x; out[4 * i + 1] = r.y; out[4 * i + 2] = r.z; out[4 * i + 3] = r.w;
}

}  // namespace dsm

// ------------------------------------------------------------------ launchers
#include "dsm_launch.h"

namespace dsm {

hipError_t launch_kmc(const KmcArgs& a, int geom, bool plain, int grid, hipStream_t st) {
  static_assert(kNumGeoms == 6, "one case per geometry below");
  switch (geom) {
    case 0: return launch_kmc_geom_0(a, plain, grid, st);
    case 1: return launch_kmc_geom_1(a, plain, grid, st);
    case 2: return launch_kmc_geom_2(a, plain, grid, st);
    case 3: return launch_kmc_geom_3(a, plain, grid, st);
    case 4: return launch_kmc_geom_4(a, plain, grid, st);
    case 5: return launch_kmc_geom_5(a, plain, grid, st);
    case kGeomPair2: return launch_kmc_geom_7(a, plain, grid, st);
    default: return hipErrorInvalidValue;
  }
}

int kmc_max_blocks_per_cu(int geom, bool plain, int lds_floats, bool ext) {
  switch (geom) {
    case 0: return kmc_occupancy_geom_0(plain, lds_floats, ext);
    case 1: return kmc_occupancy_geom_1(plain, lds_floats, ext);
    case 2: return kmc_occupancy_geom_2(plain, lds_floats, ext);
    case 3: return kmc_occupancy_geom_3(plain, lds_floats, ext);
    case 4: return kmc_occupancy_geom_4(plain, lds_floats, ext);
    case 5: return kmc_occupancy_geom_5(plain, lds_floats, ext);
    case kGeomPair2: return kmc_occupancy_geom_7(plain, lds_floats, ext);
    default: return 0;
  }
}


hipError_t launch_init_eq(const ChainArrays& ch, const EnsParams* ens, const float* tabs, const int* chains, int n,
                          hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(init_eq_kernel, dim3((n + 255) / 256), dim3(256), 0, st, ch, ens, tabs, chains, n);
  return hipGetLastError();
}

hipError_t launch_fill_u32(void* p, uint32_t v, int64_t n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(fill_u32_kernel, dim3((unsigned)((n + 255) / 256)), dim3(256), 0, st, (uint32_t*)p, v, n);
  return hipGetLastError();
}

hipError_t launch_set_ev_stop(const uint64_t* ev, uint64_t* ev_stop, int chain0, int n, uint64_t add,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(set_ev_stop_kernel, dim3((n + 255) / 256), dim3(256), 0, st, ev, ev_stop, chain0, n, add);
  return hipGetLastError();
}

hipError_t launch_sum_counters(const uint64_t* ev, const uint64_t* zsum, int chain0, int n, unsigned long long* out,
                               hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int grid = n > 256 * 4096 ? 256 : (n + 4095) / 4096;  // integer sums: the order does not matter
  hipLaunchKernelGGL(sum_counters_kernel, dim3(grid), dim3(1024), 0, st, ev, zsum, chain0, n, out);
  return hipGetLastError();
}

hipError_t launch_split_geoms(const int* src, int n, const int* Z, const int* stall, int* lists, int cap,
                              int* counts, int margin_div, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(split_geoms_kernel, dim3((n + 255) / 256), dim3(256), 0, st, src, n, Z, stall, lists, cap,
                     counts, margin_div);
  return hipGetLastError();
}

hipError_t launch_init_counts(int* cnt, int n, int n0, hipStream_t st) {
  hipLaunchKernelGGL(init_counts_kernel, dim3((n + 255) / 256), dim3(256), 0, st, cnt, n, n0);
  return hipGetLastError();
}

hipError_t launch_z_order(const int* src, int first, int n, const int* Z, const int* stall, int* scratch17, int* dst,
                          int* n_out, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int blocks = n > 64 * 1024 * 4 ? 256 : (n + 4095) / 4096 + 1;  // a few elements per thread
  hipLaunchKernelGGL(zhist_kernel, dim3(blocks), dim3(1024), 0, st, src, first, n, Z, stall, scratch17);
  hipLaunchKernelGGL(zscan_kernel, dim3(1), dim3(64), 0, st, scratch17, n_out);
  hipLaunchKernelGGL(zscatter_kernel, dim3(blocks), dim3(1024), 0, st, src, first, n, Z, stall, scratch17, dst);
  return hipGetLastError();
}

hipError_t launch_iota(int* p, int start, int n, hipStream_t st) {
  if (n <= 0) return hipSuccess;
  hipLaunchKernelGGL(iota_kernel, dim3((n + 255) / 256), dim3(256), 0, st, p, start, n);
  return hipGetLastError();
}

hipError_t launch_flow_reduce(const float* stage, int nch, int nrows, double* out,
                              hipStream_t st) {
  if (nrows <= 0) return hipSuccess;
  hipLaunchKernelGGL(flow_reduce_kernel, dim3(nrows), dim3(256), 0, st, stage, nch, nrows, out);
  return hipGetLastError();
}

hipError_t launch_corr_update(const float* stage, double* D, double* A, double* C, const int* stall, int chain0,
                              int nch, int p, int m, int s, int nrows, const int* cnt, int type, hipStream_t st) {
  if (nrows <= 0 || nch <= 0) return hipSuccess;
  CorrArgs a{stage, D, A, C, stall, chain0, nch, p, m, s, nrows, cnt, type};
  if (p == 16)
    hipLaunchKernelGGL(corr_update_reg_kernel<16>, dim3((nch + 255) / 256), dim3(256), 0, st, a);
  else
    hipLaunchKernelGGL(corr_update_kernel, dim3((nch + 255) / 256), dim3(256), 0, st, a);
  return hipGetLastError();
}

hipError_t launch_corr_reduce(const double* C, int ld, const int* stall, int chain0, int first, int count, int npts,
                              double* out_sum, int* out_healthy, hipStream_t st) {
  if (npts <= 0) return hipSuccess;
  hipLaunchKernelGGL(corr_reduce_kernel, dim3(npts), dim3(256), 0, st, C, ld, stall, chain0, first, count, npts,
                     out_sum, out_healthy);
  return hipGetLastError();
}

hipError_t launch_stage_window(const WinParam* wp, int n_ens, EnsParams* P, const int* cnt_src, int* cnt_dst,
                               int n_cnt, hipStream_t st) {
  hipLaunchKernelGGL(stage_window_kernel, dim3(1), dim3(256), 0, st, wp, n_ens, P, cnt_src, cnt_dst, n_cnt);
  return hipGetLastError();
}

hipError_t launch_chain_stress(const ChainArrays& ch, int chain0, int n, const float kappa[9], double* part,
                               hipStream_t st) {
  KappaArg K;
  for (int i = 0; i < 9; ++i) K.k[i] = kappa[i];
  hipLaunchKernelGGL(chain_stress_kernel, dim3(kStressBlocks), dim3(256), 0, st, ch, chain0, n, K, part);
  return hipGetLastError();
}

hipError_t launch_check_state(const ChainArrays& ch, int chain0, int n, double t_min, unsigned long long* out,
                              hipStream_t st) {
  if (n <= 0) return hipSuccess;
  const int grid = std::min(1024, (n + 255) / 256);
  hipLaunchKernelGGL(check_state_kernel, dim3(grid), dim3(256), 0, st, ch, chain0, n, t_min, out);
  return hipGetLastError();
}

hipError_t launch_copy_probe(const void* in, void* out, size_t n4, int grid, hipStream_t st) {
  hipLaunchKernelGGL(copy_probe_kernel, dim3(grid), dim3(256), 0, st, (const float4*)in, (float4*)out, n4);
  return hipGetLastError();
}
hipError_t launch_scan_probe(const float* in, float* out, int n, int form, hipStream_t st) {
  hipLaunchKernelGGL(scan_probe_kernel, dim3(1), dim3(64), 0, st, in, out, n, form);
  return hipGetLastError();
}
hipError_t launch_com_probe(const float* in, float* out, int n, int geom, hipStream_t st) {
  hipLaunchKernelGGL(com_probe_kernel, dim3(1), dim3(64), 0, st, in, out, n, geom);
  return hipGetLastError();
}
hipError_t launch_math_probe(const float* in, float* out, int n, int which, hipStream_t st) {
  hipLaunchKernelGGL(math_probe_kernel, dim3((n + 255) / 256), dim3(256), 0, st, in, out, n, which);
  return hipGetLastError();
}
hipError_t launch_philox_probe(const uint32_t* ck, uint32_t* out, int n, hipStream_t st) {
  hipLaunchKernelGGL(philox_probe_kernel, dim3((n + 255) / 256), dim3(256), 0, st, ck, out, n);
  return hipGetLastError();
}

}  // namespace dsm
