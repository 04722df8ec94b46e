// dsm_types.h -- data layout shared by the HIP kernels and the host engine.
//
// Chain state lives in HBM as structure-of-arrays.  Per chain (length n_chains):
//   Z, nk, ens, cap, stall, cid (Philox chain id within its ensemble),
//   local (index within its ensemble), base (offset of strand 0),
//   t (FP64 clock), ev (event counter), dt_prev (FP32, flow), next_sample, ev_stop, zsum,
//   R1 (3 floats: MSD fixed-frame position of the first entanglement).
// Per strand slot (length total capacity, chain c owns [base[c], base[c]+cap[c])):
//   Qx, Qy, Qz (FP32), N (int32), w (FP32 CD destruction rate of the link to
//   the right of the strand).
// Strand i of a chain is read by lane i%64 of slot i/64 of the wave that owns
// the chain, so a load of one slot is a contiguous 256-byte access per array.
#pragma once
#include <stdint.h>

namespace dsm {

constexpr int kNumTables = 9;  // SP SM PP PM EC ED ED2 T3 CC, each nmax+2 floats; then FR
// CD partner lengths of a polydisperse background: equiprobable quantiles, indexed by
// the top kPartnerBits bits of a Philox word
constexpr int kPartnerBits = 16;
constexpr int kPartnerK = 1 << kPartnerBits;
enum SampleMode : int { kSampleNone = 0, kSampleEq = 1, kSampleFlow = 2, kSampleMsd = 3 };

struct EnsParams {
  float beta;
  int nk;            // chain length (monodisperse) or maximum (polydisperse)
  int nmax;          // table extent
  int cd;
  int flow;
  int exp_wait;
  int sample_mode;   // SampleMode
  int nchains;
  float kappa[9];
  float c_create;
  float s_At, s_Ct, s_Dt, s_Adt, s_Bdt, s_Cdt, s_Ddt, s_taud_inv, s_one_minus_g;
  uint32_t cut_threshold;
  uint32_t key0, key1;
  int tab_off;       // float offset of SP in the table arena
  int fr_off;        // float offset of FR
  int chain0;        // first global chain index
  int partner;       // 0: CD partners have this ensemble's Nk; 1: partner tables at pd_off
  int pd_off;        // float offset of the partner tables (5K F consts, 5K D consts, K, K Nk)
  int pad0;
  double delta;      // sample / output interval
  double t_end;      // target time of the current run
  int64_t k0;        // first sample index held by the staging window
  int64_t k_end;     // samples k < k_end may be recorded in this window
  int64_t stage_off; // float offset of this ensemble's staging rows
};

// Window parameters of one ensemble in a captured pass (written by the host into
// pinned memory before each replay; stage_window_kernel copies them into EnsParams)
struct WinParam {
  int ens;
  int pad;
  double t_end;
  int64_t k0;
  int64_t k_end;
};

struct ChainArrays {
  int* Z;
  int* nk;
  int* ens;
  int* cap;
  int* stall;
  uint32_t* cid;
  int* local;
  int64_t* base;
  double* t;
  uint64_t* ev;
  float* dt_prev;
  int64_t* next_sample;
  uint64_t* ev_stop;  // jump-count target of the current fixed-count pass
  uint64_t* zsum;     // sum over the chain's jumps of Z at the jump (byte model)
  float* R1;          // [3*chain] MSD: first entanglement in the fixed frame
  float* Qx;
  float* Qy;
  float* Qz;
  int* N;
  float* w;
};

struct KmcArgs {
  ChainArrays ch;
  const EnsParams* ens;
  const float* tabs;
  const int* work;      // chain indices to process
  const int* n_work;    // number of entries in work (device memory)
  int* queue;           // persistent-form work counter
  int* overflow;        // chains that outgrew this kernel's slot capacity
  int* n_overflow;
  float* stage;         // sample staging arena
  int use_ev_stop;      // stop each chain when ev reaches ch.ev_stop[c]
  int lds_floats;       // per-wave LDS table copy size (0: tables read from global memory)
  int ext;              // an ensemble of the pass samples MSD(t), has a polydisperse CD background or flows
  int plain_jumps;      // plain form: jumps per chain per launch
};

}  // namespace dsm
