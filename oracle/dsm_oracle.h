/*
 * dsm_oracle.h -- CPU oracle for the Discrete Slip-Link Model ensemble KMC path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain C restatement of the reference
 * algorithm (ktaletsk/gpu_dsm, the common/ directory) used to check the HIP product.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may
 * load it, and only as the checker.  The product (gpu_dsm_amd) never links
 * or calls anything in this directory.
 *
 * Two modes:
 *   - mirror-FP32: performs exactly the FP32 operations of the HIP kernels,
 *     in the same order (rate formula, canonical scan tree, Philox streams,
 *     software exp/log/sincos), so integer chain state and Q are bit-exact.
 *   - FP64: the same algorithm and the same Philox streams with every
 *     floating-point quantity in double and libm transcendentals; used to
 *     state the error of the FP32 results.
 *
 * Pinning (see tests/test_oracle_*.py): Philox KAT vectors, exp/log/sincos
 * error bounds against libm double, closed-form equilibrium distributions
 * (binomial Z, uniform compositions N|Z, Gaussian Q), detailed balance of
 * every sliding-dynamics move, p_CD closed forms (normalisation, <1/tau>),
 * correlator against brute-force autocorrelation (stress and MSD forms), conservation
 * of the fixed-frame entanglement positions under every event (MSD R1 tracking),
 * stationarity of long runs.
 */
#ifndef DSM_ORACLE_H
#define DSM_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Philox counter word 3 ("purpose") values; shared convention with the HIP kernels. */
#define ORC_PURPOSE_STEP      0u   /* x0: event pick, x1: exponential wait, x2: CD lifetime */
#define ORC_PURPOSE_GAUSS     1u   /* Box-Muller -> 3 normals for a created strand */
#define ORC_PURPOSE_INIT_CUT  16u  /* equilibrium entanglement positions */
#define ORC_PURPOSE_INIT_Q    17u  /* equilibrium strand vectors */
#define ORC_PURPOSE_INIT_TAU  18u  /* equilibrium CD lifetimes */
#define ORC_PURPOSE_INIT_NK   19u  /* polydisperse chain length */
#define ORC_PURPOSE_INIT_PARTNER 20u /* CD partner quantile of each initial link (top 16 bits) */
#define ORC_PARTNER_BITS 16
#define ORC_PARTNER_K (1 << ORC_PARTNER_BITS)

/* event kinds (also the GPU's event codes) */
enum {
  ORC_EV_NONE = 0, ORC_EV_SHIFT_P = 1, ORC_EV_SHIFT_M = 2, ORC_EV_CDD = 3,
  ORC_EV_CDC = 4, ORC_EV_CDC_TAIL = 5, ORC_EV_SDC_HEAD = 6, ORC_EV_SDC_TAIL = 7,
  ORC_EV_SDD_HEAD = 8, ORC_EV_SDD_TAIL = 9, ORC_EV_STALL = 10
};

/* ---------------- random numbers ---------------- */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
float orc_u01(uint32_t x);          /* ((x>>9)+0.5)*2^-23, exact, in (0,1) */
double orc_u01_d(uint32_t x);       /* same value as double */

/* ---------------- mirror FP32 math ---------------- */
float orc_expf(float x);
float orc_logf(float x);
void orc_sincos2pif(float v, float* s, float* c);  /* sin/cos(2*pi*v), v in [0,1) */
float orc_powf(float x, float y);
void orc_boxmuller(uint32_t a, uint32_t b, float* g0, float* g1);
void orc_vec_math(int which, const float* in, float* out, int n);

/* ---------------- p_CD lifetime model (common/pcd_tau.h:26-106) ---------------- */
typedef struct {
  float g, alpha, tau_0, tau_max, tau_d;
  float c1, cm1, rcm1c1, At, Adt, Bdt, normdt;
  /* sampler constants as the device holds them (common/gpu_random.cu:76-99) */
  float s_At, s_Ct, s_Dt, s_Adt, s_Bdt, s_Cdt, s_Ddt, s_taud_inv, s_one_minus_g;
} orc_pcd;
void orc_pcd_init(orc_pcd* p, float beta, float nk);
float orc_pcd_rate_ft(const orc_pcd* p, float u);   /* SD-created / initial link: 1/tau */
float orc_pcd_rate_fdt(const orc_pcd* p, float u);  /* CD-created link: 1/tau */
double orc_pcd_rate_ft_d(const orc_pcd* p, double u);
double orc_pcd_rate_fdt_d(const orc_pcd* p, double u);

/* ---------------- ensemble parameters + FP32 tables ---------------- */
typedef struct {
  float beta;
  int nk;          /* chain length (monodisperse) or maximum (polydisperse) */
  int nmax;        /* table extent in N */
  int cd;          /* constraint dynamics on/off */
  int flow;        /* any kappa != 0 */
  int exp_wait;    /* 0: dt = 1/W (reference); 1: dt = -ln(u)/W */
  int flow_samples;/* advance records 8-component flow stress (t >= k*delta) instead of G(t) samples */
  int msd;         /* MSD(t): track R1 and record centre-of-mass samples instead of stress */
  float kappa[9];  /* xx xy xz yx yy yz zx zy zz */
  orc_pcd pcd;
  float c_create;  /* CD creation prefactor normdt/beta (common/ensemble.cu:136-150) */
  uint32_t cut_threshold; /* Bernoulli(1/(1+beta)) threshold on a 32-bit word */
  /* tables indexed by N (size nmax+2) */
  float *SP, *SM, *PP, *PM, *EC, *ED, *ED2, *T3, *CC;
  float *FR;       /* size 4*nmax+4: FR[x] = 4/x */
  int fr_size;
  /* polydisperse CD background (common/gamma.cpp, common/gpu_random.cu:62-76,212-228):
     ORC_PARTNER_K equiprobable partner-length quantiles for SD-created / initial links
     (pf) and for CD-created links (pdd, re-weighted by normdt) */
  int partner;     /* 0 none, 1 GEX, 2 the ensemble's own lengths */
  orc_pcd *pf, *pdd;
  float *pf_nk, *pd_nk;
} orc_params;
int orc_params_init(orc_params* p, float beta, int nk, int nmax, int cd, const float kappa[9], int exp_wait);
void orc_params_free(orc_params* p);

/* ---------------- FP32 mirror chain ---------------- */
typedef struct {
  int cap, Z, nk, stall;
  double t;
  uint64_t ev;
  float dt_prev;
  int64_t next_sample;
  uint32_t chain_id;
  int *N;
  float *Qx, *Qy, *Qz, *w;
  float *Q2, *Wp, *Wm, *R;   /* cached per-pair rates: pure functions of the state */
  float R1[3];               /* MSD: first entanglement in the fixed frame (common/chain.h:41) */
  int last_event;
} orc_chain;

int orc_chain_alloc(orc_chain* c, int cap);
void orc_chain_free(orc_chain* c);
void orc_chain_init_eq(orc_chain* c, const orc_params* p, const uint32_t key[2], uint32_t chain_id, int nk);
void orc_chain_refresh(orc_chain* c, const orc_params* p);   /* recompute every cached rate */
/* one KMC jump; returns event kind */
int orc_chain_event(orc_chain* c, const orc_params* p, const uint32_t key[2]);
/* canonical FP32 scan tree over positions 0..n-1 (blocks of 64 = 16 lanes x 4 slots,
   serial carry between blocks); writes inclusive sums G.  orc_tree_total is G at the
   last position of the last occupied block (zeros beyond n). */
void orc_scan_tree(const float* v, int n, float* incl);
float orc_tree_total(const float* v, int n);
/* EQ sample: (sxy, syz, sxz) */
void orc_chain_stress_sample(const orc_chain* c, const orc_params* p, float out[3]);
/* flow stress: xx yy zz xy yz zx Lpp Z */
void orc_chain_stress_full(const orc_chain* c, const orc_params* p, float out[8]);
/* MSD sample: centre of mass, sum_i (sum_{k<i} Q_k + Q_i/2) N_i / Nk + R1 */
void orc_chain_com_sample(const orc_chain* c, float out[3]);
void orc_chain_deform(orc_chain* c, const orc_params* p);

/* advance until t >= t_end (or max_events jumps); records samples k
   (k*delta < t, k < k_end) at samples + (k - k0)*stride; returns jumps done */
int64_t orc_chain_advance(orc_chain* c, const orc_params* p, const uint32_t key[2], double t_end,
                          int64_t max_events, double delta, int64_t k0, int64_t k_end, float* samples,
                          int64_t stride);

/* ---------------- FP64 chain ---------------- */
typedef struct {
  int cap, Z, nk, stall;
  double t;
  uint64_t ev;
  double dt_prev;
  int64_t next_sample;
  uint32_t chain_id;
  int *N;
  double *Qx, *Qy, *Qz, *w;
  double R1[3];
  int last_event;
} orc_chain64;
int orc_chain64_alloc(orc_chain64* c, int cap);
void orc_chain64_free(orc_chain64* c);
void orc_chain64_init_eq(orc_chain64* c, const orc_params* p, const uint32_t key[2], uint32_t chain_id, int nk);
int orc_chain64_event(orc_chain64* c, const orc_params* p, const uint32_t key[2]);
void orc_chain64_stress_sample(const orc_chain64* c, double out[3]);
void orc_chain64_com_sample(const orc_chain64* c, double out[3]);
void orc_chain64_stress_full(const orc_chain64* c, double out[8]);
int64_t orc_chain64_advance(orc_chain64* c, const orc_params* p, const uint32_t key[2], double t_end,
                            int64_t max_events, double delta, int64_t k0, int64_t k_end, double* samples,
                            int64_t stride);
/* FP64 rates for detailed-balance tests: W+ and W- of pair e */
void orc_chain64_pair_rates(const orc_chain64* c, const orc_params* p, int e, double* wp, double* wm);
double orc_chain64_end_rates(const orc_chain64* c, const orc_params* p, double out[5]);

/* ---------------- multi-tau correlator (common/eq_ensemble_kernel.cu:133-227) ---------------- */
typedef struct {
  int p, m, s, nchains;
  int type;      /* 0: stress autocorrelation, block averages; 1: MSD, (a-b)^2, first value of a block */
  int64_t nsamples;
  /* per level counters (identical for all chains: sampling is synchronous) */
  int *ins, *macc;
  int64_t *nins;
  double *D;     /* [chain][level][p][3] */
  double *A;     /* [chain][level][3] */
  double *C;     /* [level][p]: ensemble sum of (Cx+Cy+Cz) */
  double *n;     /* [level][p] */
} orc_corr;
int orc_corr_init(orc_corr* cr, int p, int m, int s, int nchains);
void orc_corr_set_type(orc_corr* cr, int type);
void orc_corr_free(orc_corr* cr);
/* add one synchronous sample row: w[chain*3 + comp] */
void orc_corr_add_row(orc_corr* cr, const double* w);
/* finalize: lags (sample units) and G values; returns number of points np */
int orc_corr_result(const orc_corr* cr, double* lag, double* g);
int orc_corr_levels(double sim_time, double delta, int p, int m);

/* ---------------- CPU baseline timing ---------------- */
/* runs n_chains chains (ids chain0..) of an equilibrium ensemble to t_end on
   n_threads threads; returns total events, writes wall seconds of the event loop */
int64_t orc_bench(const orc_params* p, const uint32_t key[2], int n_chains, uint32_t chain0,
                  double t_end, int n_threads, int fp64, double* seconds);

/* G(t) of single chains over a long horizon without storing the samples: chains chain0 ..
   chain0+n-1 of an equilibrium ensemble (per-chain Nk when nk_per_chain != NULL), each run
   from its equilibrium draw to nsamples*delta in windows of the same samples the device
   uses (a window ends at the first jump past its last sample time), every chain feeding
   its own multi-tau correlator (p, m, s levels).  out[c*np + j] = chain c's G at point j
   (np = p + (s-1)(p - p/m)); the ensemble G is the mean over chains.  fp64 selects the
   FP64 chain.  Returns the jumps made; *seconds = wall time. */
int64_t orc_gt_chains(const orc_params* p, const uint32_t key[2], int n, uint32_t chain0, const int* nk_per_chain,
                      double delta, int64_t nsamples, int64_t window, int pm, int mm, int s, int n_threads, int fp64,
                      double* out, double* seconds);

/* ---------------- ensemble API (drives many chains; used through ctypes) ---------------- */
typedef struct orc_ens orc_ens;
orc_params* orc_params_new(float beta, int nk, int nmax, int cd, const float* kappa, int exp_wait);
void orc_params_delete(orc_params* p);
void orc_params_set_flow_samples(orc_params* p, int on);
void orc_params_set_msd(orc_params* p, int on);
/* partner mode 1 (gex = a, b, mp, Mk) or 2 (nk[n] chain lengths); rebuilds c_create and
   the CC table; returns 0 on success */
int orc_params_set_partner(orc_params* p, int mode, const float* gex, const int* nk, int n);
double orc_gamma_p(double s, double x);  /* regularised lower incomplete gamma */
float orc_params_rate_ft(const orc_params* p, float u);
float orc_params_rate_fdt(const orc_params* p, float u);
/* which: 0 SP 1 SM 2 PP 3 PM 4 EC 5 ED 6 ED2 7 T3 8 CC 9 FR, 10/11 partner f_t / f_d_t sampler
   constants (5 per quantile, the device layout), 12/13 partner Nk; returns element count */
int orc_params_table(const orc_params* p, int which, float* out);
void orc_params_scalars(const orc_params* p, float* out);  /* 24 floats, see implementation */
orc_ens* orc_ens_new(const orc_params* p, uint32_t key0, uint32_t key1, uint32_t chain0, int nchains,
                     const int* nk_per_chain, int fp64);
void orc_ens_delete(orc_ens* e);
/* advance every chain; samples (float for mirror, double for fp64) laid out [k-k0][chain][ncomp],
   ncomp = 3 (G(t)) or 8 (flow_samples) */
int64_t orc_ens_advance(orc_ens* e, double t_end, int64_t max_events, double delta, int64_t k0,
                        int64_t k_end, void* samples, int n_threads);
/* state export: Z[n], N[n*cap], Q[n*cap*3], w[n*cap] (float; fp64 rounds), t[n], ev[n], stall[n] */
void orc_ens_get_r1(const orc_ens* e, float* r1);  /* [n*3]; fp64 rounds */
int orc_ens_cap(const orc_ens* e);
void orc_ens_get(const orc_ens* e, int* Z, int* N, float* Q, float* w, double* t, uint64_t* ev, int* stall);
void orc_ens_get64(const orc_ens* e, double* Q, double* w);
/* state import (mirror mode): same layout; cached rates are recomputed */
void orc_ens_set(orc_ens* e, const int* Z, const int* N, const float* Q, const float* w, const double* t,
                 const uint64_t* ev, const float* dt_prev, const int64_t* next_sample);
/* per-chain stress: mirror FP32 values (or fp64) out[n*8] as double */
void orc_ens_stress(const orc_ens* e, double* out);
/* checks */
void orc_pair_rate_f32(const orc_params* p, int Ni, int Nj, float q2i, float q2j, int endi, int endj,
                       float* wp, float* wm);
double orc_ens_cache_error(const orc_ens* e);
int orc_ens_invariants(const orc_ens* e);

#ifdef __cplusplus
}
#endif
#endif
