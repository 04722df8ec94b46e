/*
 * dsm.h -- C ABI of gpu_dsm_amd, the MI355X (gfx950) Discrete Slip-Link Model
 * ensemble KMC engine.  Plain C types only; no HIP or torch types cross it.
 *
 * Each entry point names the reference (ktaletsk/gpu_dsm) interface it replaces.
 * The reference exposes one non-reentrant, global-state ensemble per process
 * (common/ensemble.h:54-77) driven by main_cuda (common/main_cuda.cu:34); this
 * ABI is handle based so several ensembles (jobs) can share one GPU.
 *
 * Errors: every call returns DSM_OK (0) or a negative DSM_ERR_* code and never
 * calls exit(); dsm_last_error() describes the last failure of the calling thread.
 * A handle may be used by one thread at a time; separate handles are independent.
 */
#ifndef GPU_DSM_AMD_DSM_H
#define GPU_DSM_AMD_DSM_H
#include <stdbool.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define DSM_OK 0
#define DSM_ABORTED (-1)      /* run_flag cleared; reference returns -1 (common/ensemble_call_block.cu:214,353) */
#define DSM_ERR_ARG (-2)
#define DSM_ERR_DEVICE (-3)   /* reference exits with code 1 (common/cudautil.cu:33-36) */
#define DSM_ERR_MEMORY (-4)
#define DSM_ERR_IO (-5)
#define DSM_ERR_MISMATCH (-6) /* -l file holds a different N_cha; reference exit(2) (common/ensemble.cu:493-496) */
#define DSM_ERR_STATE (-7)
#define DSM_ERR_NOKERNEL (-8) /* HIP code object missing or not loadable on this device */

#define DSM_SAMPLE_NONE 0
#define DSM_SAMPLE_GT 1       /* off-diagonal stress samples for G(t) (common/eq_ensemble_kernel.cu:281-296) */
#define DSM_SAMPLE_FLOW 2     /* full stress every delta (common/ensemble_kernel.cu:504-538) */
#define DSM_SAMPLE_MSD 3      /* centre-of-mass samples for MSD(t) (common/eq_ensemble_kernel.cu:297-323) */

#define DSM_MODE_BATCHED 0    /* persistent, register-resident, Z-binned work queue */
#define DSM_MODE_PLAIN 1      /* one jump per chain per launch, state in HBM */

typedef struct dsm_engine dsm_engine;

/* One ensemble (the contents of input.dat, common/main_cuda.cu:53-68, plus the
   new knobs kept out of input.dat). */
typedef struct {
  int32_t n_chains;
  int32_t nk;                   /* Kuhn steps per chain (monodisperse) */
  const int32_t* nk_per_chain;  /* optional per-chain Nk (polydisperse); NULL = nk */
  float beta;
  int32_t cd;                   /* constraint dynamics 0/1 */
  float kappa[9];               /* velocity gradient, row-major xx xy xz yx yy yz zx zy zz */
  int32_t exp_wait;             /* 0: dt = 1/W as the reference; 1: exponential waiting times */
  uint32_t seed;                /* Philox key word 0 (the job_ID for the CLI) */
  uint32_t salt;                /* Philox key word 1 (distinguishes jobs sharing a seed) */
  int32_t sample_mode;          /* DSM_SAMPLE_* */
  double delta;                 /* sample (G) or output (flow) interval in tau_K */
  int32_t corr_p;               /* multi-tau points per level (reference 64) */
  int32_t corr_m;               /* multi-tau block factor (reference 8) */
  int32_t corr_levels;          /* levels s; 0 = derive from the run length */
  int32_t window;               /* samples per staging window; 0 = default */
  int32_t partner_mode;         /* length of the chain across a CD link (constraint-release partner):
                                   DSM_PARTNER_OWN, DSM_PARTNER_GEX or DSM_PARTNER_MELT */
  float gex[4];                 /* DSM_PARTNER_GEX: a, b, mp, Mk of polydisp.dat (common/main_cuda.cu:95-103) */
} dsm_ensemble_desc;

#define DSM_PARTNER_OWN 0     /* the chain's own Nk (reference PD_flag 0) */
#define DSM_PARTNER_GEX 1     /* GEX weight distribution (reference PD_flag 1, common/gamma.cpp:174-263) */
#define DSM_PARTNER_MELT 2    /* the ensemble's own chain lengths, weight-averaged (polydisperse melt) */

typedef struct {
  uint64_t events;      /* KMC jumps performed */
  uint64_t z_sum;       /* sum over jumps of the chain's Z (algorithmic byte model) */
  double kmc_ms;        /* device time of the KMC passes (HIP events: the sum of the launches of a
                           sequential pass; the span from split to late round of a concurrent one) */
  double first_kmc_ms;  /* device time of each pass's first launch only (the row form; the wave form
                           for a pass of few chains) */
  double total_ms;      /* host wall time of the call */
  int64_t kmc_launches;
  int64_t first_launches;
} dsm_run_stats;

/* Device selection and engine set-up; replaces checkCUDA (common/cudautil.cu:27-53)
   and the device half of gpu_init (common/ensemble.cu:118-208). */
int dsm_create(int32_t device_id, dsm_engine** out);
/* Frees the engine's device memory and streams; replaces gpu_clean
   (common/ensemble.cu:506-539) without resetting the device. */
int dsm_destroy(dsm_engine* e);
const char* dsm_last_error(void);
/* free/total HBM bytes of a device (serving admission). */
int dsm_mem_info(int32_t device_id, size_t* free_bytes, size_t* total_bytes);

/* Scheduling options of an engine (no counterpart in the reference, whose block size,
   chains per launch and sync interval are compile-time constants: common/ensemble.cu:47,
   common/textures_surfaces.h:22-24, common/ensemble_kernel.cu:41-42).  No option changes
   a result -- every chain makes the same jumps with the same random numbers under any
   setting -- only how passes are laid out on the device.  The autotuner
   (gpu_dsm_amd/autotune.py) sweeps them. */
#define DSM_OPT_PLAIN_JUMPS 1       /* plain form: jumps per chain per launch (default 1, the reference's) */
#define DSM_OPT_SMALL_PER_CU 2      /* passes of at most this many chains per CU run in the wave form (12) */
#define DSM_OPT_SPLIT_MARGIN 3      /* concurrent pass: chains within capacity/value of a wave-form geometry
                                       start in the next one (16; 0 = none) */
#define DSM_OPT_LDS_TABLE_FLOATS 4  /* rate tables go to LDS when they fit this many floats per wave (2816) */
#define DSM_OPT_CONCURRENT 5        /* geometries side by side: -1 by the pass's lengths, 0 never, 1 always */
#define DSM_OPT_BLOCKS_PER_CU 6     /* cap on persistent waves per CU (0 = the occupancy limit) */
#define DSM_OPT_WINDOW 7            /* samples per staging window of ensembles added later (256) */
#define DSM_OPT_GRAPH 8             /* 1: batched windows run as captured HIP graphs, replayed with the
                                       window's parameters in a pinned block (0) */
#define DSM_OPT_WIDE_PAIR 9         /* concurrent pass of one ensemble: chains with 128 <= Z < 240 run two
                                       per wave and those with 64 <= Z < 128 four (1), or one and two (0)
                                       (1) */
int dsm_set_option(dsm_engine* e, int32_t key, int32_t value);
int dsm_get_option(dsm_engine* e, int32_t key, int32_t* value);

/* Allocates an ensemble's chain state and tables; replaces gpu_init/init_call_block
   (common/ensemble.cu:118-208, common/ensemble_call_block.cu:55-104) and the
   p_CD set-up (common/pcd_tau.h:32-63, common/gpu_random.cu:59-104). */
int dsm_add_ensemble(dsm_engine* e, const dsm_ensemble_desc* d, int32_t* ens_id);
/* Releases an ensemble's slots for reuse by later ensembles. */
int dsm_remove_ensemble(dsm_engine* e, int32_t ens);
/* Equilibrium start; replaces host_chains_init/chain_init (common/ensemble.cu:106-115,
   common/chain.cu:131-215).  Runs on the GPU, keyed on Philox(seed, salt). */
int dsm_init_equilibrium(dsm_engine* e, int32_t ens);

/* Chain state in caller-owned host buffers, strand stride `cap` (>= max Nk):
   Z[n], N[n*cap], Q[n*cap*3], w[n*cap], t[n], ev[n], stall[n], dt_prev[n],
   next_sample[n]; any pointer may be NULL.  Replaces get_chains_from_device /
   get_chain_to_device_call_block (common/ensemble.cu:231-236,
   common/ensemble_call_block.cu:402-450). */
int dsm_state_cap(dsm_engine* e, int32_t ens, int32_t* cap);
int dsm_get_state(dsm_engine* e, int32_t ens, int32_t cap, int32_t* Z, int32_t* N, float* Q, float* w,
                  double* t, uint64_t* ev, int32_t* stall, float* dt_prev, int64_t* next_sample);
int dsm_set_state(dsm_engine* e, int32_t ens, int32_t cap, const int32_t* Z, const int32_t* N, const float* Q,
                  const float* w, const double* t, const uint64_t* ev, const int32_t* stall,
                  const float* dt_prev, const int64_t* next_sample);
/* Size, strand capacity and sample mode (DSM_SAMPLE_*) of an ensemble; any pointer may be NULL. */
int dsm_ensemble_info(dsm_engine* e, int32_t ens, int32_t* n_chains, int32_t* cap, int32_t* sample_mode);
/* The -s / -l binary files (int32 N_cha; per chain a 16-byte head {int32 Z, float time,
   float unused, int32 stall}, Z float4 {Qx Qy Qz N}, Z float CD rates): dsm_save writes the
   ensemble's state with times relative to t_ref; dsm_load reads a file in the reference's
   layout into the ensemble (jump counters restart at 0, sampling restarts at the loaded
   clocks) and returns DSM_ERR_MISMATCH for another chain count or DSM_ERR_IO for a
   missing, truncated or invalid file (dsm_load_error() says which).  Replace save_to_file
   / load_from_file (common/ensemble.cu:342-354,486-504, common/chain.cu:236-250), which
   exit(2) on a mismatch. */
int dsm_save(dsm_engine* e, int32_t ens, const char* path, double t_ref);
int dsm_load(dsm_engine* e, int32_t ens, const char* path);
const char* dsm_load_error(void);
/* Ensemble stress of the current state: out[0..7] = averages over the chains that have
   not stalled of xx yy zz xy yz zx Lpp Z (per chain sigma_ab = -sum 3 Q_a Q_b / N, FP64
   from the stored FP32 state, the deformation pending from the last jump applied);
   *healthy (may be NULL) = their count.  Replaces calc_stress (common/ensemble.cu:210-221)
   over stress_calc (common/ensemble_kernel.cu:504-538). */
int dsm_stress(dsm_engine* e, int32_t ens, double* out, int32_t* healthy);
/* Integrity scan of an ensemble's state on the device (no counterpart in the reference):
   out[0..6] = chains with Z outside 1..Nk, with a strand N < 1, with sum N != Nk, with a
   chain-end strand vector != 0, with a non-finite strand vector or CD rate, stalled
   chains, and chains whose clock is below t_min. */
int dsm_check_state(dsm_engine* e, int32_t ens, double t_min, int64_t* out);
/* Chain state in the caller's own device memory, structure of arrays: per chain Z,
   stall flag, base (index of strand 0 in the strand arrays) and dt_prev (the waiting time
   of its last jump, whose deformation is still pending); per strand slot Qx, Qy, Qz, N. */
typedef struct {
  const int32_t* Z;
  const int32_t* stall;
  const int64_t* base;
  const float* dt_prev;
  const float* Qx;
  const float* Qy;
  const float* Qz;
  const int32_t* N;
} dsm_device_chains;
/* dsm_stress for chain state in the caller's device memory, allocating nothing: no
   counterpart in the reference; for callers that cannot allocate (code under HIP graph
   capture, tight serving loops).  Same kernel and result as dsm_stress (out[0..7] =
   averages over the chains that have not stalled of xx yy zz xy yz zx Lpp Z, kappa the
   velocity gradient for the pending deformation, NULL = none).  Its partial sums live in
   a fixed device array of 256 x 9 doubles (18 KiB) on each GPU the library runs on, kept
   until the process ends; every chain count fits it (the grid is fixed), and calls that
   overlap on one GPU share it and race.  Synchronous on the device's null stream. */
int dsm_stress_noalloc(int32_t device_id, const dsm_device_chains* chains, int32_t n_chains, const float* kappa,
                       double* out, int32_t* healthy);
/* MSD mode: R1[n*3], each chain's first entanglement in the fixed frame, the chain_head
   R1 field (common/chain.h:41) that the reference keeps on the device (d_R1,
   common/ensemble_call_block.cu:75,406,422).  Zero after dsm_init_equilibrium. */
int dsm_get_r1(dsm_engine* e, int32_t ens, float* r1);

/* Advances every chain of the ensemble until its clock reaches t_end, recording
   samples and updating the correlator / flow stress on the way.  run_flag (may be
   NULL) is polled after each sampling window (its samples consumed): when cleared the
   call returns DSM_ABORTED with every chain at that window's end, a state a later call
   continues exactly; progress (may be NULL) receives 0..100.
   Replaces gpu_time_step/time_step_call_block (common/ensemble.cu:223-229,
   common/ensemble_call_block.cu:106-222) and the G(t) loop of gpu_Gt_PCS /
   EQ_time_step_call_block (common/ensemble.cu:289-339,
   common/ensemble_call_block.cu:224-379). */
int dsm_advance(dsm_engine* e, int32_t ens, double t_end, int32_t mode, volatile const int32_t* run_flag,
                volatile int32_t* progress, dsm_run_stats* stats);
/* Several ensembles (different Nk, beta, kappa, targets) in one batched launch
   sequence: the persistent kernels take chains of all of them from one queue. */
int dsm_advance_many(dsm_engine* e, const int32_t* ens_ids, const double* t_end, int32_t count, int32_t mode,
                     volatile const int32_t* run_flag, dsm_run_stats* stats);
/* Exactly n_events jumps per chain (no time target); for parity tests and benches. */
int dsm_advance_events(dsm_engine* e, int32_t ens, int64_t n_events, int32_t mode, dsm_run_stats* stats);

/* Flow stress rows recorded so far: out[row*9 + q] = sum over healthy chains of
   xx yy zz xy yz zx Lpp Z (q < 8) and the healthy-chain count (q = 8), row r is
   time (r+1)*delta.  Replaces calc_stress/calc_stress_call_block
   (common/ensemble.cu:210-221, common/ensemble_call_block.cu:452-505). */
int dsm_flow_stress(dsm_engine* e, int32_t ens, int64_t first_row, int64_t max_rows, double* out,
                    int64_t* n_rows);
/* G(t): lags in samples and G values (np points, reference order); replaces
   correlator::calc and the accumulation in gpu_Gt_PCS (common/correlator.cu:146-187,
   common/ensemble.cu:309-321). */
int dsm_correlator(dsm_engine* e, int32_t ens, int32_t max_pts, double* lag, double* g, int32_t* npts);
/* The same over chains first .. first+count-1 of the ensemble only (the chains' own
   mean; error bars over chain groups). */
int dsm_correlator_range(dsm_engine* e, int32_t ens, int32_t first, int32_t count, int32_t max_pts, double* lag,
                         double* g, int32_t* npts);
/* The samples of the last staging window: out[(k-k0)*n*3 + chain*3 + comp]. */
int dsm_last_samples(dsm_engine* e, int32_t ens, float* out, int64_t max_floats, int64_t* k0, int64_t* nrows);

/* Inspection of the FP32 tables / p_CD constants an ensemble runs with
   (which: 0 SP 1 SM 2 PP 3 PM 4 EC 5 ED 6 ED2 7 T3 8 CC 9 FR; with a polydisperse CD
   background, 10/11 the f_t / f_d_t sampler constants of the 65536 partner quantiles
   (5 floats each) and 12/13 the partner Nk of those quantiles). */
int dsm_tables(dsm_engine* e, int32_t ens, int32_t which, float* out, int32_t* n);
int dsm_ensemble_scalars(dsm_engine* e, int32_t ens, float* out24);

/* Building-block probes (run one kernel on the device; no counterpart in the
   reference): the canonical FP32 scan tree of the KMC kernels (form 1: one chain of
   n <= 1024 positions, out = n inclusive sums + total; form 0: four chains of
   n <= 64 positions in the row form, in = 4*n values, out = 4*(n+1); form 2: two
   chains of n <= 128 positions in the pair form, in = 2*n, out = 2*(n+1); form 3: two
   chains of n <= 256 positions in the pair form with two register groups; form 4: four
   chains of n <= 128 positions in the row form with two register groups; form 5: one
   chain of n <= 512 positions in the wave form with two register groups, as form 1),
   exp/log/sin/cos (which 0..3), Philox (6 words in, 4 out). */
int dsm_probe_scan(int32_t device_id, const float* in, float* out, int32_t n, int32_t form);
int dsm_probe_math(int32_t device_id, const float* in, float* out, int32_t n, int32_t which);
int dsm_probe_philox(int32_t device_id, const uint32_t* ctr_key, uint32_t* out, int32_t n);
/* MSD centre-of-mass sample of one chain state (relative to R1) in KMC geometry `geom`
   (0 row form, n <= 64; 1 pair form, n <= 128; 2/3/4 wave form, n <= 256/512/1024;
   7 pair form with two register groups, n <= 256; 8 row form with two register groups,
   n <= 128):
   in[4i..4i+3] = Qx Qy Qz N
   of strand i, in[4n] = Nk; out = 3 floats. */
int dsm_probe_com(int32_t device_id, const float* in, float* out, int32_t n, int32_t geom);
/* Plain device-to-device copy of n_bytes (multiple of 16), timed over `reps` launches
   after one warm-up: the achievable-bandwidth and byte-counter reference (no
   counterpart in the reference). */
int dsm_probe_copy(int32_t device_id, int64_t n_bytes, int32_t reps, double* ms_per_copy);

/* Options of the drop-in driver that input.dat cannot carry. */
typedef struct {
  int32_t mode;         /* DSM_MODE_* */
  int32_t corr_p;       /* default 64 */
  int32_t corr_m;       /* default 8 */
  int32_t window;       /* 0 = default */
  int32_t exp_wait;     /* default 0 */
  int32_t quiet;        /* suppress per-line stdout echo */
  const char* workdir;  /* directory holding input.dat and receiving outputs; NULL = cwd */
} dsm_main_options;

/* Drop-in for main_cuda (common/main_cuda.cu:34): reads input.dat, runs the
   flow / G(t) calculation, writes tau/G/MSD[_ID].dat, distr_*.dat and the -s file.
   Returns 0, -1 when run_flag is cleared, or a DSM_ERR_* code. */
int dsm_main(volatile const bool* run_flag, int job_ID, const char* savefile, const char* loadfile,
             int device_ID, bool distr, volatile int* progress_bar);
int dsm_main_ex(volatile const bool* run_flag, int job_ID, const char* savefile, const char* loadfile,
                int device_ID, bool distr, volatile int* progress_bar, const dsm_main_options* opt);

/* Serving (no counterpart in the reference, whose driver runs one job per process):
   a seeded stream of simulation requests with Poisson arrivals is admitted into one
   engine while the requests' estimated device memory fits hbm_fraction of the free
   HBM, every admitted request advances by `window` tau_K per scheduling round
   (one dsm_advance_many pass), and a request completes when its horizon is reached
   and its result (G(t), or flow stress rows when sheared) has been read back. */
typedef struct {
  int32_t n_requests;
  double arrival_rate;       /* requests per second (wall clock) */
  uint64_t seed;             /* request stream; also Philox key word 0 (salt = request index) */
  int32_t nk_min, nk_max, nk_step;  /* Nk uniform over nk_min, nk_min+step, ..., <= nk_max (<= 2047) */
  int32_t n_beta;            /* 1..4 */
  float beta[4];             /* beta uniform over the first n_beta entries */
  int32_t nch_min, nch_max;  /* chains per request, log-uniform */
  double flow_fraction;      /* share of requests under shear (others: equilibrium G(t)) */
  double rate_min, rate_max; /* shear rate kappa_xy, log-uniform, 1/tau_K */
  double horizon;            /* tau_K simulated per request */
  double window;             /* tau_K per scheduling round */
  int32_t cd;                /* constraint dynamics */
  double hbm_fraction;       /* share of the free device memory admission may fill */
  double progress_s;         /* > 0: one status line on stderr every progress_s seconds */
  int32_t graph;             /* 1: scheduling rounds run as captured graphs (DSM_OPT_GRAPH) */
} dsm_serve_config;

typedef struct {
  int32_t completed;
  int32_t max_active;        /* most requests resident at once */
  int32_t rounds;
  double makespan_s;         /* first arrival window start to last completion */
  double p50_s, p99_s, mean_s, max_s;  /* arrival-to-result latency */
  uint64_t jumps;
  double jumps_per_s;        /* aggregate over the makespan */
  int64_t peak_bytes;        /* largest admitted memory estimate */
  double admit_s, advance_s, fetch_s;  /* host wall time in admission (allocation and
                                          equilibrium draw), KMC rounds, result retrieval */
} dsm_serve_report;

/* Runs the stream on one device; latency_s (optional) receives n_requests values. */
int dsm_serve(int32_t device_id, const dsm_serve_config* cfg, dsm_serve_report* report, double* latency_s);
/* As dsm_serve; summary (optional) receives DSM_SERVE_SUMMARY doubles per request, read
   with its result: [0] = 1 (G(t)) or 2 (flow), [1] = points or rows of the result; G(t):
   [2..17] = G at lags 0..15 samples; flow: [2..10] = the last output row (sums over the
   healthy chains of xx yy zz xy yz zx Lpp Z, then their count). */
#define DSM_SERVE_SUMMARY 18
int dsm_serve_ex(int32_t device_id, const dsm_serve_config* cfg, dsm_serve_report* report, double* latency_s,
                 double* summary);
/* The request stream alone (host only): arrival times, Nk, beta, chains, shear rate
   (0 = equilibrium) and the admission memory estimate of each request. */
int dsm_serve_plan(const dsm_serve_config* cfg, int32_t max_requests, double* arrival_s, int32_t* nk, float* beta,
                   int32_t* n_chains, double* shear_rate, int64_t* bytes);

#ifdef __cplusplus
}
#endif
#endif
