/*
 * dsm_oracle.c -- CPU oracle (test infrastructure) for the DSM ensemble KMC path.
 * See dsm_oracle.h for scope and the rule that the product never uses this file.
 *
 * Reference behaviour restated here (ktaletsk/gpu_dsm):
 *   rates ............ common/ensemble_kernel.cu:149-181, common/eq_ensemble_kernel.cu:85-126
 *   end rates ........ common/eq_ensemble_kernel.cu:340-394
 *   clock + pick ..... common/eq_ensemble_kernel.cu:395-419 (dt = 1/W, pr = W*u, linear scan)
 *   event effects .... common/eq_ensemble_kernel.cu:420-662, common/ensemble_kernel.cu:290-501
 *   deformation ...... common/ensemble_kernel.cu:105-110,135-146 (pinned in tests/test_oracle_flow.py:
 *                      exact affine map from the clock's dt sequence, linear response,
 *                      Green-Kubo, second-order N1)
 *   stress ........... common/ensemble_kernel.cu:504-538, common/eq_ensemble_kernel.cu:281-296
 *                      (full tensor, Lpp and Z pinned by closed forms at rest)
 *   p_CD ............. common/pcd_tau.h:32-82, common/gpu_random.cu:76-99,143-148
 *   CD prefactor ..... common/ensemble.cu:136-150
 *   init ............. common/chain.cu:38-215 (same distributions; re-keyed on Philox)
 *   correlator ....... common/eq_ensemble_kernel.cu:133-227, common/correlator.cu:104-187
 *
 * Deliberate, documented differences from the reference (identical in the HIP kernels):
 *   - randomness is Philox4x32-10 keyed by (seed, salt), counter (event, chain, purpose);
 *   - the equilibrium draw uses independent Bernoulli(1/(1+beta)) cuts at every inner
 *     Kuhn-step boundary, which gives exactly the reference's joint law of Z
 *     (binomial) and N|Z (uniform compositions);
 *   - chain ends are identified by index (0 and Z-1) rather than by |Q| == 0;
 *   - rate totals are FP32 sums in a fixed tree order (the reference sums serially):
 *     see the canonical scan tree below;
 *   - CD creation at the tail is gated by the CD flag in flow mode too.
 */
#define _GNU_SOURCE
#include "dsm_oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <pthread.h>
#include <time.h>

/* ------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11)                   */
/* ------------------------------------------------------------------ */
void orc_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
  uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
  uint32_t k0 = key[0], k1 = key[1];
  for (int r = 0; r < 10; ++r) {
    if (r) { k0 += 0x9E3779B9u; k1 += 0xBB67AE85u; }
    uint64_t p0 = (uint64_t)0xD2511F53u * c0;
    uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
    uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    uint32_t n0 = hi1 ^ c1 ^ k0, n1 = lo1, n2 = hi0 ^ c3 ^ k1, n3 = lo0;
    c0 = n0; c1 = n1; c2 = n2; c3 = n3;
  }
  out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* 23 random bits: (k + 0.5) * 2^-23 is exact in FP32, so u lies in [2^-24, 1 - 2^-24] */
float orc_u01(uint32_t x) { return ((float)(x >> 9) + 0.5f) * 0x1p-23f; }
double orc_u01_d(uint32_t x) { return ((double)(x >> 9) + 0.5) * 0x1p-23; }

/* ------------------------------------------------------------------ */
/* FP32 math, mirrored bit-for-bit by the HIP kernels                   */
/* exp: Cody-Waite reduction by ln2 + degree-7 Taylor polynomial        */
/* log: exponent split + atanh series 2(s + s^3/3 + ... + s^11/11)      */
/* sin/cos(2 pi v): exact quadrant reduction on v + Taylor polynomials  */
/* ------------------------------------------------------------------ */
static inline float as_float(uint32_t u) { float f; memcpy(&f, &u, 4); return f; }
static inline uint32_t as_uint(float f) { uint32_t u; memcpy(&u, &f, 4); return u; }

float orc_expf(float x) {
  if (x != x) return x;
  if (x > 88.0f) return INFINITY;
  if (x < -87.0f) return 0.0f;
  const float n = rintf(x * 0x1.715476p+0f);
  float r = fmaf(-n, 0x1.62e4p-1f, x);
  r = fmaf(-n, 0x1.7f7d1cp-20f, r);
  float q = 0x1.a01a02p-13f;
  q = fmaf(q, r, 0x1.6c16c2p-10f);
  q = fmaf(q, r, 0x1.111112p-7f);
  q = fmaf(q, r, 0x1.555556p-5f);
  q = fmaf(q, r, 0x1.555556p-3f);
  q = fmaf(q, r, 0x1p-1f);
  q = fmaf(q, r, 1.0f);
  q = fmaf(q, r, 1.0f);
  const int ni = (int)n;
  return q * as_float((uint32_t)(ni + 127) << 23);
}

float orc_logf(float x) {
  uint32_t ix = as_uint(x);
  int e = (int)(ix >> 23) - 127;
  uint32_t mb = (ix & 0x7fffffu) | 0x3f800000u;
  if (mb > 0x3fb504f3u) { mb -= 0x00800000u; e += 1; }
  const float m = as_float(mb);
  const float f = m - 1.0f;
  const float s = f / (2.0f + f);
  const float z = s * s;
  float q = 0x1.745d18p-3f;
  q = fmaf(z, q, 0x1.c71c72p-3f);
  q = fmaf(z, q, 0x1.24924ap-2f);
  q = fmaf(z, q, 0x1.99999ap-2f);
  q = fmaf(z, q, 0x1.555556p-1f);
  const float lnm = fmaf(s, z * q, s + s);
  const float fe = (float)e;
  return fmaf(fe, 0x1.62e4p-1f, fmaf(fe, 0x1.7f7d1cp-20f, lnm));
}

void orc_sincos2pif(float v, float* sn, float* cs) {
  const float t = v - rintf(v);
  const float qf = rintf(4.0f * t);
  const float f = t - 0.25f * qf;
  const float z = f * f;
  float ps = 0x1.507834p+5f;
  ps = fmaf(z, ps, -0x1.32d2ccp+6f);
  ps = fmaf(z, ps, 0x1.466bc6p+6f);
  ps = fmaf(z, ps, -0x1.4abbcep+5f);
  ps = fmaf(z, ps, 0x1.921fb6p+2f);
  const float s = f * ps;
  float pc = -0x1.a6d1f2p+4f;
  pc = fmaf(z, pc, 0x1.e1f506p+5f);
  pc = fmaf(z, pc, -0x1.55d3c8p+6f);
  pc = fmaf(z, pc, 0x1.03c1f0p+6f);
  pc = fmaf(z, pc, -0x1.3bd3ccp+4f);
  const float c = fmaf(z, pc, 1.0f);
  switch (((int)qf) & 3) {
    case 0: *sn = s; *cs = c; break;
    case 1: *sn = c; *cs = -s; break;
    case 2: *sn = -s; *cs = -c; break;
    default: *sn = -c; *cs = s; break;
  }
}

float orc_powf(float x, float y) { return orc_expf(y * orc_logf(x)); }

void orc_boxmuller(uint32_t a, uint32_t b, float* g0, float* g1) {
  const float r = sqrtf(-2.0f * orc_logf(orc_u01(a)));
  float s, c;
  orc_sincos2pif(orc_u01(b), &s, &c);
  *g0 = r * c;
  *g1 = r * s;
}

static void boxmuller_d(uint32_t a, uint32_t b, double* g0, double* g1) {
  const double r = sqrt(-2.0 * log(orc_u01_d(a)));
  const double th = 2.0 * M_PI * orc_u01_d(b);
  *g0 = r * cos(th);
  *g1 = r * sin(th);
}

/* ------------------------------------------------------------------ */
/* p_CD: common/pcd_tau.h:32-63 (FP32 exactly as the reference host)  */
/* ------------------------------------------------------------------ */
void orc_pcd_init(orc_pcd* p, float Be, float Nk) {
  double z = (Nk + Be) / (Be + 1.0);
  p->g = 0.667f;
  if (Be != 1.0f) {
    p->alpha = (0.053f * logf(Be) + 0.31f) * powf((float)z, -0.012f * logf(Be) - 0.024f);
    p->tau_0 = 0.285f * powf(Be + 2.0f, 0.515f);
    p->tau_max = ((Nk < 2) ? p->tau_0 : 0.025f * powf(Be + 2.0f, 2.6f) * powf((float)z, 2.83f));
    p->tau_d = ((Nk < 2) ? p->tau_0 : 0.036f * powf(Be + 2.0f, 3.07f) * powf((float)(z - 1.0f), 3.02f));
  } else {
    p->alpha = 0.267096f - 0.375571f * expf(-0.0838237f * Nk);
    p->tau_0 = 0.460277f + 0.298913f * expf(-0.0705314f * Nk);
    p->tau_max = ((Nk < 4) ? p->tau_0 : 0.0156137f * powf(Nk, 3.18849f));
    p->tau_d = ((Nk < 4) ? p->tau_0 : 0.0740131f * powf(Nk, 3.18363f));
  }
  p->c1 = powf(p->tau_max, p->alpha) - powf(p->tau_0, p->alpha);
  p->cm1 = powf(p->tau_max, p->alpha - 1.0f) - powf(p->tau_0, p->alpha - 1.0f);
  p->rcm1c1 = ((p->c1 == 0.0f) ? (p->alpha - 1.0f) / p->alpha / p->tau_0 : (p->cm1 / p->c1));
  p->At = (1.0f - p->g);
  p->Adt = p->At * p->alpha / (p->alpha - 1.0f);
  p->Bdt = p->Adt * p->rcm1c1;
  p->normdt = p->Bdt + p->g / p->tau_d;
  /* device sampler constants, common/gpu_random.cu:76-99 */
  p->s_taud_inv = 1.0f / p->tau_d;
  p->s_At = 1.0f * p->c1 / p->At;
  p->s_Dt = powf(p->tau_0, p->alpha);
  p->s_Ct = -1.0f / p->alpha;
  p->s_Adt = p->normdt * p->c1 / p->Adt;
  p->s_Bdt = p->Bdt / p->normdt;
  p->s_Cdt = -1.0f / (p->alpha - 1.0f);
  p->s_Ddt = powf(p->tau_0, p->alpha - 1.0f);
  p->s_one_minus_g = 1.0f - p->g;
}

/* common/gpu_random.cu:146-148 (returns the rate 1/tau) */
float orc_pcd_rate_ft(const orc_pcd* p, float u) {
  return u < p->s_one_minus_g ? orc_powf(fmaf(u, p->s_At, p->s_Dt), p->s_Ct) : p->s_taud_inv;
}
/* common/gpu_random.cu:143-145 */
float orc_pcd_rate_fdt(const orc_pcd* p, float u) {
  return u < p->s_Bdt ? orc_powf(fmaf(u, p->s_Adt, p->s_Ddt), p->s_Cdt) : p->s_taud_inv;
}
double orc_pcd_rate_ft_d(const orc_pcd* p, double u) {
  return u < (double)p->s_one_minus_g ? pow(u * p->s_At + p->s_Dt, (double)p->s_Ct) : 1.0 / (double)p->tau_d;
}
double orc_pcd_rate_fdt_d(const orc_pcd* p, double u) {
  return u < (double)p->s_Bdt ? pow(u * p->s_Adt + p->s_Ddt, (double)p->s_Cdt) : 1.0 / (double)p->tau_d;
}

/* ------------------------------------------------------------------ */
/* ensemble parameters and FP32 tables                                 */
/* ------------------------------------------------------------------ */
static double pow075(double x) { return sqrt(x) * sqrt(sqrt(x)); }

int orc_params_init(orc_params* p, float beta, int nk, int nmax, int cd, const float kappa[9], int exp_wait) {
  memset(p, 0, sizeof(*p));
  if (nmax < nk) nmax = nk;
  p->beta = beta; p->nk = nk; p->nmax = nmax; p->cd = cd ? 1 : 0; p->exp_wait = exp_wait ? 1 : 0;
  p->flow = 0;
  p->flow_samples = 0;
  for (int i = 0; i < 9; ++i) {
    p->kappa[i] = kappa ? kappa[i] : 0.0f;
    if (p->kappa[i] != 0.0f) p->flow = 1;
  }
  orc_pcd_init(&p->pcd, beta, (float)nk);
  p->c_create = p->pcd.normdt / beta;
  {
    double pc = 1.0 / (1.0 + (double)beta);
    double T = floor(pc * 4294967296.0);
    if (T > 4294967295.0) T = 4294967295.0;
    p->cut_threshold = (uint32_t)T;
  }
  const int n = nmax + 2;
  p->SP = calloc(n, 4); p->SM = calloc(n, 4); p->PP = calloc(n, 4); p->PM = calloc(n, 4);
  p->EC = calloc(n, 4); p->ED = calloc(n, 4); p->ED2 = calloc(n, 4); p->T3 = calloc(n, 4);
  p->CC = calloc(n, 4);
  p->fr_size = 4 * nmax + 4;
  p->FR = calloc(p->fr_size, 4);
  if (!p->SP || !p->FR) return -1;
  for (int N = 1; N < n; ++N) {
    const float fN = (float)N;
    p->SP[N] = 0.75f / (fN * (float)(N + 1));
    p->SM[N] = N >= 2 ? 0.75f / (fN * (float)(N - 1)) : 0.0f;
    p->PP[N] = (float)pow075((double)N / (double)(N + 1));
    p->PM[N] = N >= 2 ? (float)pow075((double)N / (double)(N - 1)) : 0.0f;
    p->EC[N] = 2.0f / (beta * (fN + 0.5f));
    p->ED[N] = 1.0f / (0.75f + 0.5f * fN);
    p->ED2[N] = 1.0f / (0.75f + (fN + 0.25f));
    p->T3[N] = 3.0f / fN;
    p->CC[N] = p->cd ? p->c_create * (float)(N - 1) : 0.0f;
  }
  for (int x = 1; x < p->fr_size; ++x) p->FR[x] = 4.0f / (float)x;
  return 0;
}

void orc_params_free(orc_params* p) {
  free(p->SP); free(p->SM); free(p->PP); free(p->PM); free(p->EC); free(p->ED);
  free(p->ED2); free(p->T3); free(p->CC); free(p->FR);
  free(p->pf); free(p->pdd); free(p->pf_nk); free(p->pd_nk);
  memset(p, 0, sizeof(*p));
}

/* ------------------------------------------------------------------ */
/* Polydisperse CD background                                          */
/* ------------------------------------------------------------------ */
/* P(s, x) = gamma(s, x) / Gamma(s): power series for x < s + 1, else 1 - Q with Q from
   Legendre's continued fraction evaluated by the modified Lentz method. */
double orc_gamma_p(double s, double x) {
  if (x <= 0.0) return 0.0;
  const double lead = -x + s * log(x) - lgamma(s);
  if (x < s + 1.0) {
    double term = 1.0 / s, sum = term;
    for (int n = 1; n < 100000; ++n) {
      term *= x / (s + n);
      sum += term;
      if (fabs(term) < fabs(sum) * 1e-17) break;
    }
    return sum * exp(lead);
  }
  const double tiny = 1e-300;
  double b = x + 1.0 - s, cc = 1.0 / tiny, d = 1.0 / b, h = d;
  for (int i = 1; i < 100000; ++i) {
    const double an = -i * (i - s);
    b += 2.0;
    d = an * d + b;
    if (fabs(d) < tiny) d = tiny;
    cc = b + an / cc;
    if (fabs(cc) < tiny) cc = tiny;
    d = 1.0 / d;
    const double del = d * cc;
    h *= del;
    if (fabs(del - 1.0) < 1e-16) break;
  }
  return 1.0 - exp(lead) * h;
}

static int cmp_int(const void* a, const void* b) {
  const int x = *(const int*)a, y = *(const int*)b;
  return (x > y) - (x < y);
}

/* Quantile k of the partner weight distribution is the length at probability (k+1/2)/K:
   mode 1 inverts the GEX weight fraction W(m) = P((a+1)/b, m^b) (common/gamma.cpp:174-231)
   by bisection and scales Nk = m*mp/Mk; mode 2 weights each distinct chain length by
   itself times its count.  The f_d_t table takes the same entries re-weighted by their
   mean destruction rate normdt (common/gamma.cpp:236-262); c_create becomes the mean
   normdt over the entries divided by beta (common/ensemble.cu:137-146). */
int orc_params_set_partner(orc_params* p, int mode, const float* gex, const int* nk, int n) {
  const int K = ORC_PARTNER_K;
  free(p->pf); free(p->pdd); free(p->pf_nk); free(p->pd_nk);
  p->pf = calloc(K, sizeof(orc_pcd)); p->pdd = calloc(K, sizeof(orc_pcd));
  p->pf_nk = calloc(K, 4); p->pd_nk = calloc(K, 4);
  if (mode == 1) {
    const double a = gex[0], b = gex[1], mp = gex[2], Mk = gex[3];
    if (!(a > -1.0) || !(b > 0.0) || !(mp > 0.0) || !(Mk > 0.0)) return -1;
    const double s = (a + 1.0) / b;
    for (int k = 0; k < K; ++k) {
      const double y = (k + 0.5) / K;
      double lo = 0.0, hi = 1.0;
      while (orc_gamma_p(s, hi) < y) hi *= 2.0;
      for (int it = 0; it < 64; ++it) {
        const double mid = 0.5 * (lo + hi);
        if (orc_gamma_p(s, mid) < y) lo = mid; else hi = mid;
      }
      const double m = pow(0.5 * (lo + hi), 1.0 / b);
      p->pf_nk[k] = (float)(m * mp / Mk);
    }
  } else if (mode == 2) {
    int* v = malloc(sizeof(int) * (size_t)n);
    memcpy(v, nk, sizeof(int) * (size_t)n);
    qsort(v, (size_t)n, sizeof(int), cmp_int);
    double tot = 0.0;
    for (int i = 0; i < n; ++i) tot += (double)v[i];
    /* distinct values in ascending order with weight value * count */
    int i = 0;
    double cum = 0.0;
    int cur = v[0];
    while (i < n && v[i] == cur) { cum += (double)v[i]; ++i; }
    for (int k = 0; k < K; ++k) {
      const double y = (k + 0.5) / K * tot;
      while (cum < y && i < n) {
        cur = v[i];
        while (i < n && v[i] == cur) { cum += (double)v[i]; ++i; }
      }
      p->pf_nk[k] = (float)cur;
    }
    free(v);
  } else {
    return -1;
  }
  double* cum = malloc(sizeof(double) * K);
  double tot = 0.0;
  for (int k = 0; k < K; ++k) {
    orc_pcd_init(&p->pf[k], p->beta, p->pf_nk[k]);
    tot += (double)p->pf[k].normdt;
    cum[k] = tot;
  }
  int i = 0;
  for (int j = 0; j < K; ++j) {
    const double y = (j + 0.5) / K * tot;
    while (cum[i] < y && i + 1 < K) ++i;
    p->pdd[j] = p->pf[i];
    p->pd_nk[j] = p->pf_nk[i];
  }
  free(cum);
  p->partner = mode;
  p->c_create = (float)(tot / K) / p->beta;
  for (int N = 1; N < p->nmax + 2; ++N) p->CC[N] = p->cd ? p->c_create * (float)(N - 1) : 0.0f;
  return 0;
}

/* sampler constants for a new link: the ensemble's own, or those of the partner quantile
   in the top bits of Philox word w */
static const orc_pcd* pcd_ft(const orc_params* p, uint32_t w) {
  return p->partner ? &p->pf[w >> (32 - ORC_PARTNER_BITS)] : &p->pcd;
}
static const orc_pcd* pcd_fdt(const orc_params* p, uint32_t w) {
  return p->partner ? &p->pdd[w >> (32 - ORC_PARTNER_BITS)] : &p->pcd;
}

/* ------------------------------------------------------------------ */
/* Canonical FP32 scan tree over positions 0..n-1 (zeros beyond n), the */
/* summation order the HIP kernels use bit for bit.  Position           */
/* p = 64b + 4l + k: block b of 64, lane l of 16, slot k of 4.          */
/*   lane:  serial prefix s_k over the lane's 4 slots;                  */
/*   block: Hillis-Steele inclusive scan H of the 16 lane totals        */
/*          (strides 1,2,4,8; a source before lane 0 adds 0), exclusive */
/*          prefix X_l = H_{l-1} (X_0 = 0), block-local I = X_l + s_k;  */
/*   chain: G = I in block 0, G = C_b + I in block b >= 1, with         */
/*          C_1 = I(63 of block 0), C_{b+1} = C_b + I(63 of block b).   */
/* The total of n positions is G at the last position of the last       */
/* occupied block (64*ceil(n/64) - 1).                                  */
/* ------------------------------------------------------------------ */
static void block_scan(const float v[64], float I[64]) {
  float s[64], H[16], t[16];
  for (int l = 0; l < 16; ++l) {
    float acc = v[4 * l];
    s[4 * l] = acc;
    for (int k = 1; k < 4; ++k) { acc = acc + v[4 * l + k]; s[4 * l + k] = acc; }
    H[l] = acc;
  }
  for (int d = 1; d <= 8; d <<= 1) {
    for (int l = 0; l < 16; ++l) t[l] = l >= d ? H[l - d] : 0.0f;
    for (int l = 0; l < 16; ++l) H[l] = H[l] + t[l];
  }
  for (int l = 0; l < 16; ++l) {
    const float X = l == 0 ? 0.0f : H[l - 1];
    for (int k = 0; k < 4; ++k) I[4 * l + k] = X + s[4 * l + k];
  }
}

/* writes G for positions 0..n-1; returns the total */
static float scan_tree_total(const float* v, int n, float* incl) {
  float buf[64], I[64];
  float carry = 0.0f;
  for (int b = 0; b * 64 < n; ++b) {
    for (int q = 0; q < 64; ++q) {
      const int i = b * 64 + q;
      buf[q] = i < n ? v[i] : 0.0f;
    }
    block_scan(buf, I);
    for (int q = 0; q < 64; ++q) {
      const int i = b * 64 + q;
      if (i < n && incl) incl[i] = b == 0 ? I[q] : carry + I[q];
    }
    carry = b == 0 ? I[63] : carry + I[63];
  }
  return carry;
}

void orc_scan_tree(const float* v, int n, float* incl) { (void)scan_tree_total(v, n, incl); }

float orc_tree_total(const float* v, int n) { return n > 0 ? scan_tree_total(v, n, NULL) : 0.0f; }

static float tree_total(const float* v, int n, float* scratch) {
  (void)scratch;
  return orc_tree_total(v, n);
}

/* ------------------------------------------------------------------ */
/* chains (FP32 mirror)                                                */
/* ------------------------------------------------------------------ */
int orc_chain_alloc(orc_chain* c, int cap) {
  memset(c, 0, sizeof(*c));
  c->cap = cap;
  c->N = calloc(cap + 2, sizeof(int));
  c->Qx = calloc(cap + 2, 4); c->Qy = calloc(cap + 2, 4); c->Qz = calloc(cap + 2, 4);
  c->w = calloc(cap + 2, 4); c->Q2 = calloc(cap + 2, 4);
  c->Wp = calloc(cap + 2, 4); c->Wm = calloc(cap + 2, 4); c->R = calloc(cap + 2, 4);
  return (c->N && c->R) ? 0 : -1;
}

void orc_chain_free(orc_chain* c) {
  free(c->N); free(c->Qx); free(c->Qy); free(c->Qz); free(c->w); free(c->Q2);
  free(c->Wp); free(c->Wm); free(c->R);
  memset(c, 0, sizeof(*c));
}

/* Equilibrium draw (distributions of common/chain.cu:38-215).  Kuhn-step
   boundary i (1 <= i < nk) carries an entanglement with probability
   1/(1+beta), independently: Z-1 ~ Binomial(nk-1, 1/(1+beta)) and, given Z,
   the cut set is a uniform (Z-1)-subset, i.e. N|Z is a uniform composition. */
static int init_strands(int* N, const orc_params* p, const uint32_t key[2], uint32_t chain_id, int nk) {
  uint32_t ctr[4] = {0, 0, chain_id, ORC_PURPOSE_INIT_CUT}, x[4] = {0, 0, 0, 0};
  int Z = 0, len = 0;
  for (int i = 1; i <= nk; ++i) {
    ++len;
    if (i == nk) { N[Z++] = len; break; }
    const int wi = i - 1;
    if ((wi & 3) == 0) { ctr[0] = (uint32_t)(wi >> 2); orc_philox4x32_10(ctr, key, x); }
    if (x[wi & 3] < p->cut_threshold) { N[Z++] = len; len = 0; }
  }
  return Z;
}

static inline void q2_of(orc_chain* c, int i) {
  c->Q2[i] = fmaf(c->Qz[i], c->Qz[i], fmaf(c->Qy[i], c->Qy[i], c->Qx[i] * c->Qx[i]));
}

void orc_chain_init_eq(orc_chain* c, const orc_params* p, const uint32_t key[2], uint32_t chain_id, int nk) {
  c->nk = nk; c->chain_id = chain_id; c->t = 0.0; c->ev = 0; c->dt_prev = 0.0f; c->stall = 0;
  c->next_sample = p->flow ? 1 : 0;  /* flow output k is at time k*delta, k >= 1 */
  c->last_event = 0;
  c->R1[0] = c->R1[1] = c->R1[2] = 0.0f;  /* common/chain.cu:207 */
  for (int i = 0; i < c->cap + 2; ++i) {
    c->N[i] = 0; c->Qx[i] = c->Qy[i] = c->Qz[i] = c->w[i] = 0.0f;
  }
  c->Z = init_strands(c->N, p, key, chain_id, nk);
  for (int k = 1; k < c->Z - 1; ++k) {
    uint32_t ctr[4] = {(uint32_t)k, 0, chain_id, ORC_PURPOSE_INIT_Q}, x[4];
    orc_philox4x32_10(ctr, key, x);
    float g0, g1, g2, g3;
    orc_boxmuller(x[0], x[1], &g0, &g1);
    orc_boxmuller(x[2], x[3], &g2, &g3);
    const float sig = sqrtf((float)c->N[k] / 3.0f);
    c->Qx[k] = g0 * sig; c->Qy[k] = g1 * sig; c->Qz[k] = g2 * sig;
  }
  if (p->cd) {
    uint32_t x[4] = {0, 0, 0, 0}, y[4] = {0, 0, 0, 0};
    for (int k = 0; k < c->Z - 1; ++k) {
      if ((k & 3) == 0) {
        uint32_t ctr[4] = {(uint32_t)(k >> 2), 0, chain_id, ORC_PURPOSE_INIT_TAU};
        orc_philox4x32_10(ctr, key, x);
        uint32_t cty[4] = {(uint32_t)(k >> 2), 0, chain_id, ORC_PURPOSE_INIT_PARTNER};
        if (p->partner) orc_philox4x32_10(cty, key, y);
      }
      c->w[k] = orc_pcd_rate_ft(pcd_ft(p, y[k & 3]), orc_u01(x[k & 3]));
    }
  }
  orc_chain_refresh(c, p);
}

/* pair rates of link e (strands i=e, j=e+1); common/eq_ensemble_kernel.cu:85-126 */
static void pair_rate(orc_chain* c, const orc_params* p, int e) {
  const int Z = c->Z;
  const int i = e, j = e + 1;
  const int Ni = c->N[i], Nj = c->N[j];
  const int endi = (i == 0), endj = (j == Z - 1);
  const float q2i = c->Q2[i], q2j = c->Q2[j];
  float wp = 0.0f, wm = 0.0f;
  if (Nj > 1) {
    const float a = q2i * p->SP[Ni];
    const float b = q2j * p->SM[Nj];
    const int x = (endi ? 4 * Ni + 1 : 2 * Ni) + (endj ? 4 * Nj - 1 : 2 * Nj);
    const float pf = (p->FR[x] * (endi ? 1.0f : p->PP[Ni])) * (endj ? 1.0f : p->PM[Nj]);
    wp = pf * orc_expf(a - b);
  }
  if (Ni > 1) {
    const float a = q2i * p->SM[Ni];
    const float b = q2j * p->SP[Nj];
    const int x = (endi ? 4 * Ni - 1 : 2 * Ni) + (endj ? 4 * Nj + 1 : 2 * Nj);
    const float pf = (p->FR[x] * (endi ? 1.0f : p->PM[Ni])) * (endj ? 1.0f : p->PP[Nj]);
    wm = pf * orc_expf(b - a);
  }
  c->Wp[e] = wp;
  c->Wm[e] = wm;
  c->R[e] = p->cd ? (wp + wm) + (c->w[e] + p->CC[Ni]) : (wp + wm);
}

/* end-event rates; common/eq_ensemble_kernel.cu:340-392.  out: cdcz sdc1 sdcz sdd1 sddz */
static float end_rates(const orc_chain* c, const orc_params* p, float out[5]) {
  const int Z = c->Z;
  float cdcz = 0.0f, sdc1 = 0.0f, sdcz = 0.0f, sdd1 = 0.0f, sddz = 0.0f;
  if (Z == 1) {
    cdcz = p->CC[c->N[0]];
    sdc1 = 1.0f / (p->beta * (float)c->nk);
    sdcz = sdc1;
  } else {
    const int N0 = c->N[0], N1 = c->N[1], Nz = c->N[Z - 1], Nz2 = c->N[Z - 2];
    if (N0 == 1) sdd1 = (Z == 2) ? p->ED2[N1] : p->ED[N1];
    else sdc1 = p->EC[N0];
    if (Nz == 1) sddz = (Z == 2) ? p->ED2[Nz2] : p->ED[Nz2];
    else sdcz = p->EC[Nz];
    cdcz = p->CC[Nz];
  }
  out[0] = cdcz; out[1] = sdc1; out[2] = sdcz; out[3] = sdd1; out[4] = sddz;
  return (((cdcz + sdc1) + sdcz) + sdd1) + sddz;
}

void orc_chain_refresh(orc_chain* c, const orc_params* p) {
  for (int i = 0; i < c->Z; ++i) q2_of(c, i);
  for (int e = 0; e < c->Z - 1; ++e) pair_rate(c, p, e);
}

static void refresh_range(orc_chain* c, const orc_params* p, int lo, int hi) {
  if (lo < 0) lo = 0;
  if (hi > c->Z - 2) hi = c->Z - 2;
  for (int e = lo; e <= hi; ++e) pair_rate(c, p, e);
}

/* move strand/link records: dst index <- src index */
static inline void mv(orc_chain* c, int dst, int src) {
  c->N[dst] = c->N[src]; c->Qx[dst] = c->Qx[src]; c->Qy[dst] = c->Qy[src]; c->Qz[dst] = c->Qz[src];
  c->w[dst] = c->w[src]; c->Q2[dst] = c->Q2[src];
  c->Wp[dst] = c->Wp[src]; c->Wm[dst] = c->Wm[src]; c->R[dst] = c->R[src];
}
static void open_gap(orc_chain* c, int at) {   /* indices >= at move up by one */
  for (int i = c->Z; i > at; --i) mv(c, i, i - 1);
}
static void close_gap(orc_chain* c, int at) {  /* index at is removed */
  for (int i = at; i < c->Z - 1; ++i) mv(c, i, i + 1);
}
static inline void set_strand(orc_chain* c, int i, int N, float x, float y, float z) {
  c->N[i] = N; c->Qx[i] = x; c->Qy[i] = y; c->Qz[i] = z; q2_of(c, i);
}

static int pick_sub(const float* s, int n, float* r) {
  float rr = *r;
  for (int m = 0; m < n; ++m) {
    if (s[m] > 0.0f && rr < s[m]) { *r = rr; return m; }
    rr = rr - s[m];
  }
  *r = rr;
  return -1;
}

static int split_newn(float r, int N, float wc) {
  int newn = (int)floorf(0.5f + (r * (float)(N - 2)) / wc) + 1;
  if (newn < 1) newn = 1;
  if (newn > N - 1) newn = N - 1;
  return newn;
}

int orc_chain_event(orc_chain* c, const orc_params* p, const uint32_t key[2]) {
  if (c->stall) return ORC_EV_STALL;
  const int Z = c->Z;
  float pos[4096 + 64], incl[4096 + 64], endr[5];
  float* posv = pos; float* inclv = incl;
  float* heap = NULL;
  if (Z > 4096) { heap = malloc(sizeof(float) * 2 * (size_t)(Z + 64)); posv = heap; inclv = heap + Z + 64; }
  for (int e = 0; e < Z - 1; ++e) posv[e] = c->R[e];
  posv[Z - 1] = end_rates(c, p, endr);
  const float W = scan_tree_total(posv, Z, inclv);
  uint32_t ctr[4] = {(uint32_t)c->ev, (uint32_t)(c->ev >> 32), c->chain_id, ORC_PURPOSE_STEP}, x[4];
  orc_philox4x32_10(ctr, key, x);
  const float dt = p->exp_wait ? (-orc_logf(orc_u01(x[1]))) / W : 1.0f / W;
  int stall = 0;
  if (dt == 0.0f) stall = 1;
  if (isnan(dt)) stall = 2;
  if (isinf(dt)) stall = 3;
  if (stall) { c->stall = stall; free(heap); return ORC_EV_STALL; }
  c->t += (double)dt;
  c->dt_prev = dt;
  const float pr = W * orc_u01(x[0]);
  int ps = -1;
  for (int q = 0; q < Z; ++q) if (inclv[q] > pr) { ps = q; break; }
  int kind = ORC_EV_NONE;
  if (ps >= 0) {
    float r = ps == 0 ? pr : pr - inclv[ps - 1];
    free(heap); heap = NULL;
    const float u_tau = orc_u01(x[2]);
    float g[4];
    int have_g = 0;
#define NEED_G() do { if (!have_g) { uint32_t gc[4] = {ctr[0], ctr[1], c->chain_id, ORC_PURPOSE_GAUSS}, gx[4]; \
      orc_philox4x32_10(gc, key, gx); orc_boxmuller(gx[0], gx[1], &g[0], &g[1]); orc_boxmuller(gx[2], gx[3], &g[2], &g[3]); have_g = 1; } } while (0)
    if (ps < Z - 1) {
      const int e = ps;
      const float s[4] = {p->cd ? c->w[e] : 0.0f, c->Wp[e], c->Wm[e], p->CC[c->N[e]]};
      const int m = pick_sub(s, 4, &r);
      if (m == 1) {          /* shift + : one Kuhn step from j into i */
        c->N[e] += 1; c->N[e + 1] -= 1;
        kind = ORC_EV_SHIFT_P;
        refresh_range(c, p, e - 1, e + 1);
      } else if (m == 2) {   /* shift - */
        c->N[e] -= 1; c->N[e + 1] += 1;
        kind = ORC_EV_SHIFT_M;
        refresh_range(c, p, e - 1, e + 1);
      } else if (m == 0) {   /* CD destruction of link e: merge strands e, e+1 */
        const int Nm = c->N[e] + c->N[e + 1];
        if (p->msd && e == 0) {  /* the second entanglement becomes the first (eq_ensemble_kernel.cu:447-455) */
          c->R1[0] = c->R1[0] + c->Qx[1]; c->R1[1] = c->R1[1] + c->Qy[1]; c->R1[2] = c->R1[2] + c->Qz[1];
        }
        float mx = 0.0f, my = 0.0f, mz = 0.0f;
        if (!(e == 0 || e == Z - 2)) {
          mx = c->Qx[e] + c->Qx[e + 1]; my = c->Qy[e] + c->Qy[e + 1]; mz = c->Qz[e] + c->Qz[e + 1];
        }
        /* strands: e <- merged, i > e <- old i+1; links: i >= e <- old i+1 */
        const float w_next = c->w[e + 1];
        close_gap(c, e + 1);
        c->w[e] = w_next;
        c->Z = Z - 1;
        set_strand(c, e, Nm, mx, my, mz);
        kind = ORC_EV_CDD;
        refresh_range(c, p, e - 1, e);
      } else if (m == 3) {   /* CD creation inside strand e */
        NEED_G();
        const int N = c->N[e];
        const int newn = split_newn(r, N, s[3]);
        const float wnew = orc_pcd_rate_fdt(pcd_fdt(p, x[3]), u_tau);
        if (e == 0) {
          const float sig = sqrtf((float)(N - newn) / 3.0f);
          open_gap(c, 0);
          c->Z = Z + 1;
          set_strand(c, 0, newn, 0.0f, 0.0f, 0.0f);
          c->w[0] = wnew;
          set_strand(c, 1, N - newn, g[0] * sig, g[1] * sig, g[2] * sig);
          if (p->msd) {  /* new first entanglement, one strand before the old (eq_ensemble_kernel.cu:531-539) */
            c->R1[0] = c->R1[0] - c->Qx[1]; c->R1[1] = c->R1[1] - c->Qy[1]; c->R1[2] = c->R1[2] - c->Qz[1];
          }
          refresh_range(c, p, 0, 1);
        } else {
          const float sig = sqrtf(((float)newn * (float)(N - newn)) / (3.0f * (float)N));
          const float ratio = (float)newn / (float)N;
          const float qx = c->Qx[e], qy = c->Qy[e], qz = c->Qz[e];
          const float nx = (g[0] * sig) + (qx * ratio);
          const float ny = (g[1] * sig) + (qy * ratio);
          const float nz = (g[2] * sig) + (qz * ratio);
          open_gap(c, e);
          c->Z = Z + 1;
          set_strand(c, e, newn, nx, ny, nz);
          c->w[e] = wnew;
          set_strand(c, e + 1, N - newn, qx - nx, qy - ny, qz - nz);
          refresh_range(c, p, e - 1, e + 1);
        }
        kind = ORC_EV_CDC;
      }
    } else {
      const int m = pick_sub(endr, 5, &r);
      if (m == 0) {          /* CD creation in the tail strand */
        NEED_G();
        const int N = c->N[Z - 1];
        const int newn = split_newn(r, N, endr[0]);
        const float wnew = orc_pcd_rate_fdt(pcd_fdt(p, x[3]), u_tau);
        const float sig = (Z == 1) ? 0.0f : sqrtf((float)newn / 3.0f);
        c->Z = Z + 1;
        if (Z == 1) set_strand(c, Z - 1, newn, 0.0f, 0.0f, 0.0f);
        else set_strand(c, Z - 1, newn, g[0] * sig, g[1] * sig, g[2] * sig);
        c->w[Z - 1] = wnew;
        set_strand(c, Z, N - newn, 0.0f, 0.0f, 0.0f);
        c->w[Z] = 0.0f;
        refresh_range(c, p, Z - 2, Z - 1);
        kind = ORC_EV_CDC_TAIL;
      } else if (m == 1) {   /* SD creation at the head */
        NEED_G();
        const int N0 = c->N[0];
        const float wnew = orc_pcd_rate_ft(pcd_ft(p, x[3]), u_tau);
        const float sig = (Z == 1) ? 0.0f : sqrtf((float)(N0 - 1) / 3.0f);
        open_gap(c, 0);
        c->Z = Z + 1;
        set_strand(c, 0, 1, 0.0f, 0.0f, 0.0f);
        c->w[0] = p->cd ? wnew : 0.0f;
        if (Z == 1) set_strand(c, 1, N0 - 1, 0.0f, 0.0f, 0.0f);
        else set_strand(c, 1, N0 - 1, g[0] * sig, g[1] * sig, g[2] * sig);
        if (p->msd) {  /* eq_ensemble_kernel.cu:606-614 */
          c->R1[0] = c->R1[0] - c->Qx[1]; c->R1[1] = c->R1[1] - c->Qy[1]; c->R1[2] = c->R1[2] - c->Qz[1];
        }
        refresh_range(c, p, 0, 1);
        kind = ORC_EV_SDC_HEAD;
      } else if (m == 2) {   /* SD creation at the tail */
        NEED_G();
        const int Nz = c->N[Z - 1];
        const float wnew = orc_pcd_rate_ft(pcd_ft(p, x[3]), u_tau);
        const float sig = (Z == 1) ? 0.0f : sqrtf((float)(Nz - 1) / 3.0f);
        c->Z = Z + 1;
        if (Z == 1) set_strand(c, Z - 1, Nz - 1, 0.0f, 0.0f, 0.0f);
        else set_strand(c, Z - 1, Nz - 1, g[0] * sig, g[1] * sig, g[2] * sig);
        c->w[Z - 1] = p->cd ? wnew : 0.0f;
        set_strand(c, Z, 1, 0.0f, 0.0f, 0.0f);
        c->w[Z] = 0.0f;
        refresh_range(c, p, Z - 2, Z - 1);
        kind = ORC_EV_SDC_TAIL;
      } else if (m == 3) {   /* SD destruction at the head: strands 0,1 merge into an end */
        const int N1 = c->N[1];
        if (p->msd) {  /* eq_ensemble_kernel.cu:645-653 */
          c->R1[0] = c->R1[0] + c->Qx[1]; c->R1[1] = c->R1[1] + c->Qy[1]; c->R1[2] = c->R1[2] + c->Qz[1];
        }
        close_gap(c, 0);
        c->Z = Z - 1;
        set_strand(c, 0, N1 + 1, 0.0f, 0.0f, 0.0f);
        refresh_range(c, p, 0, 0);
        kind = ORC_EV_SDD_HEAD;
      } else if (m == 4) {   /* SD destruction at the tail */
        const int Nz2 = c->N[Z - 2];
        c->Z = Z - 1;
        set_strand(c, Z - 2, Nz2 + 1, 0.0f, 0.0f, 0.0f);
        c->w[Z - 2] = 0.0f;
        refresh_range(c, p, Z - 3, Z - 3);
        kind = ORC_EV_SDD_TAIL;
      }
    }
#undef NEED_G
  }
  free(heap);
  c->ev += 1;
  c->last_event = kind;
  return kind;
}

/* common/ensemble_kernel.cu:105-110: explicit Euler step Q <- Q + dt*kappa*Q */
void orc_chain_deform(orc_chain* c, const orc_params* p) {
  const float dt = c->dt_prev;
  float dk[9];
  for (int i = 0; i < 9; ++i) dk[i] = dt * p->kappa[i];
  for (int i = 0; i < c->Z; ++i) {
    const float x = c->Qx[i], y = c->Qy[i], z = c->Qz[i];
    c->Qx[i] = fmaf(dk[2], z, fmaf(dk[1], y, fmaf(dk[0], x, x)));
    c->Qy[i] = fmaf(dk[5], z, fmaf(dk[4], y, fmaf(dk[3], x, y)));
    c->Qz[i] = fmaf(dk[8], z, fmaf(dk[7], y, fmaf(dk[6], x, z)));
  }
}

void orc_chain_stress_sample(const orc_chain* c, const orc_params* p, float out[3]) {
  const int Z = c->Z;
  float* a = malloc(sizeof(float) * 4 * (size_t)(Z + 64));
  float* b = a + Z + 64;
  float* d = b + Z + 64;
  float* s = d + Z + 64;
  for (int i = 0; i < Z; ++i) {
    const float t3 = p->T3[c->N[i]];
    a[i] = (c->Qx[i] * c->Qy[i]) * t3;
    b[i] = (c->Qy[i] * c->Qz[i]) * t3;
    d[i] = (c->Qx[i] * c->Qz[i]) * t3;
  }
  out[0] = -tree_total(a, Z, s);
  out[1] = -tree_total(b, Z, s);
  out[2] = -tree_total(d, Z, s);
  free(a);
}

void orc_chain_stress_full(const orc_chain* c, const orc_params* p, float out[8]) {
  const int Z = c->Z;
  float* v = malloc(sizeof(float) * 2 * (size_t)(Z + 64));
  float* s = v + Z + 64;
  for (int comp = 0; comp < 7; ++comp) {
    for (int i = 0; i < Z; ++i) {
      const float x = c->Qx[i], y = c->Qy[i], z = c->Qz[i], t3 = p->T3[c->N[i]];
      float val;
      switch (comp) {
        case 0: val = (x * x) * t3; break;
        case 1: val = (y * y) * t3; break;
        case 2: val = (z * z) * t3; break;
        case 3: val = (x * y) * t3; break;
        case 4: val = (y * z) * t3; break;
        case 5: val = (x * z) * t3; break;
        default: val = sqrtf(fmaf(z, z, fmaf(y, y, x * x))); break;
      }
      v[i] = val;
    }
    const float tot = tree_total(v, Z, s);
    out[comp] = comp < 6 ? -tot : tot;
  }
  out[7] = (float)Z;
  free(v);
}

/* Centre of mass (common/eq_ensemble_kernel.cu:297-323): strand i sits at its midpoint
   P_i + Q_i/2 with P_i the position of its left end relative to R1, weighted N_i/Nk.
   FP32 in the canonical order: P_i is the canonical inclusive scan of Q at i-1 (0 at
   i = 0), the weighted terms (P_i + Q_i*0.5)*N_i/Nk are summed by the canonical tree,
   then R1 is added. */
void orc_chain_com_sample(const orc_chain* c, float out[3]) {
  const int Z = c->Z;
  const float nk = (float)c->nk;
  float* q = malloc(sizeof(float) * 3 * (size_t)(Z + 64));
  float* incl = q + Z + 64;
  float* v = incl + Z + 64;
  const float* src[3] = {c->Qx, c->Qy, c->Qz};
  for (int comp = 0; comp < 3; ++comp) {
    for (int i = 0; i < Z; ++i) q[i] = src[comp][i];
    scan_tree_total(q, Z, incl);
    for (int i = 0; i < Z; ++i) {
      const float P = i > 0 ? incl[i - 1] : 0.0f;
      v[i] = ((P + q[i] * 0.5f) * (float)c->N[i]) / nk;
    }
    out[comp] = orc_tree_total(v, Z) + c->R1[comp];
  }
  free(q);
}

int64_t orc_chain_advance(orc_chain* c, const orc_params* p, const uint32_t key[2], double t_end,
                          int64_t max_events, double delta, int64_t k0, int64_t k_end, float* samples,
                          int64_t stride) {
  int64_t n = 0;
  for (;;) {
    if (p->flow) {
      if (c->dt_prev != 0.0f) { orc_chain_deform(c, p); orc_chain_refresh(c, p); }
      c->dt_prev = 0.0f;
    }
    if (samples && !p->flow_samples) {
      /* G(t) sample k: state after the first jump past k*delta (common/eq_ensemble_kernel.cu:256-331) */
      while (c->next_sample < k_end && c->t > (double)c->next_sample * delta) {
        float sv[3];
        if (p->msd) orc_chain_com_sample(c, sv);
        else orc_chain_stress_sample(c, p, sv);
        if (c->next_sample >= k0) {  /* rows k0..k_end-1 only, as the kernels */
          float* o = samples + (c->next_sample - k0) * stride;
          o[0] = sv[0]; o[1] = sv[1]; o[2] = sv[2];
        }
        c->next_sample++;
      }
    } else if (samples) {
      /* flow output k at time k*delta: chain reached when t >= k*delta (common/ensemble_kernel.cu:202-206) */
      while (c->next_sample < k_end && c->t >= (double)c->next_sample * delta) {
        float sv[8];
        orc_chain_stress_full(c, p, sv);
        if (c->next_sample >= k0) {
          float* o = samples + (c->next_sample - k0) * stride;
          for (int q = 0; q < 8; ++q) o[q] = sv[q];
        }
        c->next_sample++;
      }
    }
    if (c->stall || c->t >= t_end || n >= max_events) break;
    orc_chain_event(c, p, key);
    ++n;
  }
  return n;
}

/* ------------------------------------------------------------------ */
/* FP64 chain: same algorithm, same Philox words, double arithmetic    */
/* ------------------------------------------------------------------ */
int orc_chain64_alloc(orc_chain64* c, int cap) {
  memset(c, 0, sizeof(*c));
  c->cap = cap;
  c->N = calloc(cap + 2, sizeof(int));
  c->Qx = calloc(cap + 2, 8); c->Qy = calloc(cap + 2, 8); c->Qz = calloc(cap + 2, 8);
  c->w = calloc(cap + 2, 8);
  return (c->N && c->w) ? 0 : -1;
}
void orc_chain64_free(orc_chain64* c) {
  free(c->N); free(c->Qx); free(c->Qy); free(c->Qz); free(c->w);
  memset(c, 0, sizeof(*c));
}

void orc_chain64_init_eq(orc_chain64* c, const orc_params* p, const uint32_t key[2], uint32_t chain_id, int nk) {
  c->nk = nk; c->chain_id = chain_id; c->t = 0.0; c->ev = 0; c->dt_prev = 0.0; c->stall = 0;
  c->next_sample = p->flow ? 1 : 0;
  c->last_event = 0;
  c->R1[0] = c->R1[1] = c->R1[2] = 0.0;
  for (int i = 0; i < c->cap + 2; ++i) { c->N[i] = 0; c->Qx[i] = c->Qy[i] = c->Qz[i] = c->w[i] = 0.0; }
  c->Z = init_strands(c->N, p, key, chain_id, nk);
  for (int k = 1; k < c->Z - 1; ++k) {
    uint32_t ctr[4] = {(uint32_t)k, 0, chain_id, ORC_PURPOSE_INIT_Q}, x[4];
    orc_philox4x32_10(ctr, key, x);
    double g0, g1, g2, g3;
    boxmuller_d(x[0], x[1], &g0, &g1);
    boxmuller_d(x[2], x[3], &g2, &g3);
    const double sig = sqrt((double)c->N[k] / 3.0);
    c->Qx[k] = g0 * sig; c->Qy[k] = g1 * sig; c->Qz[k] = g2 * sig;
  }
  if (p->cd) {
    uint32_t x[4] = {0, 0, 0, 0}, y[4] = {0, 0, 0, 0};
    for (int k = 0; k < c->Z - 1; ++k) {
      if ((k & 3) == 0) {
        uint32_t ctr[4] = {(uint32_t)(k >> 2), 0, chain_id, ORC_PURPOSE_INIT_TAU};
        orc_philox4x32_10(ctr, key, x);
        uint32_t cty[4] = {(uint32_t)(k >> 2), 0, chain_id, ORC_PURPOSE_INIT_PARTNER};
        if (p->partner) orc_philox4x32_10(cty, key, y);
      }
      c->w[k] = orc_pcd_rate_ft_d(pcd_ft(p, y[k & 3]), orc_u01_d(x[k & 3]));
    }
  }
}

/* reference rate formulas in double (common/eq_ensemble_kernel.cu:97-124) */
void orc_chain64_pair_rates(const orc_chain64* c, const orc_params* p, int e, double* wp, double* wm) {
  const int Z = c->Z, i = e, j = e + 1;
  const double Ni = c->N[i], Nj = c->N[j];
  const int endi = (i == 0), endj = (j == Z - 1);
  const double Q = c->Qx[i] * c->Qx[i] + c->Qy[i] * c->Qy[i] + c->Qz[i] * c->Qz[i];
  const double Q2 = c->Qx[j] * c->Qx[j] + c->Qy[j] * c->Qy[j] + c->Qz[j] * c->Qz[j];
  *wp = 0.0; *wm = 0.0;
  if (Nj > 1.0) {
    const double sig1 = 0.75 / (Ni * (Ni + 1.0)), sig2 = 0.75 / (Nj * (Nj - 1.0));
    const double pf1 = endi ? 1.0 : Ni / (Ni + 1.0), pf2 = endj ? 1.0 : Nj / (Nj - 1.0);
    const double f1 = endi ? 2.0 * Ni + 0.5 : Ni, f2 = endj ? 2.0 * Nj - 0.5 : Nj;
    *wp = 2.0 / (f1 + f2) * pow(pf1 * pf2, 0.75) * exp(Q * sig1 - Q2 * sig2);
  }
  if (Ni > 1.0) {
    const double sig1 = 0.75 / (Ni * (Ni - 1.0)), sig2 = 0.75 / (Nj * (Nj + 1.0));
    const double pf1 = endi ? 1.0 : Ni / (Ni - 1.0), pf2 = endj ? 1.0 : Nj / (Nj + 1.0);
    const double f1 = endi ? 2.0 * Ni - 0.5 : Ni, f2 = endj ? 2.0 * Nj + 0.5 : Nj;
    *wm = 2.0 / (f1 + f2) * pow(pf1 * pf2, 0.75) * exp(-Q * sig1 + Q2 * sig2);
  }
}

double orc_chain64_end_rates(const orc_chain64* c, const orc_params* p, double out[5]) {
  const int Z = c->Z;
  const double beta = p->beta, cc = p->cd ? (double)p->c_create : 0.0;
  double cdcz = 0, sdc1 = 0, sdcz = 0, sdd1 = 0, sddz = 0;
  if (Z == 1) {
    cdcz = cc * (c->N[0] - 1);
    sdc1 = 1.0 / (beta * c->nk);
    sdcz = sdc1;
  } else {
    const double N0 = c->N[0], N1 = c->N[1], Nz = c->N[Z - 1], Nz2 = c->N[Z - 2];
    if (N0 == 1.0) sdd1 = 1.0 / (0.75 + ((Z == 2) ? N1 + 0.25 : 0.5 * N1));
    else sdc1 = 2.0 / (beta * (N0 + 0.5));
    if (Nz == 1.0) sddz = 1.0 / (0.75 + ((Z == 2) ? Nz2 + 0.25 : 0.5 * Nz2));
    else sdcz = 2.0 / (beta * (Nz + 0.5));
    cdcz = cc * (Nz - 1.0);
  }
  out[0] = cdcz; out[1] = sdc1; out[2] = sdcz; out[3] = sdd1; out[4] = sddz;
  return cdcz + sdc1 + sdcz + sdd1 + sddz;
}

static int split_newn_d(double r, int N, double wc) {
  int newn = (int)floor(0.5 + r * (double)(N - 2) / wc) + 1;
  if (newn < 1) newn = 1;
  if (newn > N - 1) newn = N - 1;
  return newn;
}

static void mv64(orc_chain64* c, int dst, int src) {
  c->N[dst] = c->N[src]; c->Qx[dst] = c->Qx[src]; c->Qy[dst] = c->Qy[src]; c->Qz[dst] = c->Qz[src];
  c->w[dst] = c->w[src];
}
static void set64(orc_chain64* c, int i, int N, double x, double y, double z) {
  c->N[i] = N; c->Qx[i] = x; c->Qy[i] = y; c->Qz[i] = z;
}

int orc_chain64_event(orc_chain64* c, const orc_params* p, const uint32_t key[2]) {
  if (c->stall) return ORC_EV_STALL;
  const int Z = c->Z;
  const double cc = p->cd ? (double)p->c_create : 0.0;
  double W = 0.0, endr[5];
  for (int e = 0; e < Z - 1; ++e) {
    double wp, wm;
    orc_chain64_pair_rates(c, p, e, &wp, &wm);
    W += wp + wm + (p->cd ? c->w[e] + cc * (c->N[e] - 1) : 0.0);
  }
  W += orc_chain64_end_rates(c, p, endr);
  uint32_t ctr[4] = {(uint32_t)c->ev, (uint32_t)(c->ev >> 32), c->chain_id, ORC_PURPOSE_STEP}, x[4];
  orc_philox4x32_10(ctr, key, x);
  const double dt = p->exp_wait ? -log(orc_u01_d(x[1])) / W : 1.0 / W;
  if (!(dt > 0.0) || isinf(dt)) { c->stall = isnan(dt) ? 2 : (dt == 0.0 ? 1 : 3); return ORC_EV_STALL; }
  c->t += dt;
  c->dt_prev = dt;
  double pr = W * orc_u01_d(x[0]);
  const double u_tau = orc_u01_d(x[2]);
  double g[4];
  int have_g = 0;
#define NEED_G64() do { if (!have_g) { uint32_t gc[4] = {ctr[0], ctr[1], c->chain_id, ORC_PURPOSE_GAUSS}, gx[4]; \
    orc_philox4x32_10(gc, key, gx); boxmuller_d(gx[0], gx[1], &g[0], &g[1]); boxmuller_d(gx[2], gx[3], &g[2], &g[3]); have_g = 1; } } while (0)
  int kind = ORC_EV_NONE;
  for (int e = 0; e < Z - 1 && kind == ORC_EV_NONE; ++e) {
    double wp, wm;
    orc_chain64_pair_rates(c, p, e, &wp, &wm);
    const double wd = p->cd ? c->w[e] : 0.0, wc = cc * (c->N[e] - 1);
    const double tot = wd + wp + wm + wc;
    if (pr >= tot) { pr -= tot; continue; }
    if (pr < wd) {
      const int Nm = c->N[e] + c->N[e + 1];
      if (p->msd && e == 0) { c->R1[0] += c->Qx[1]; c->R1[1] += c->Qy[1]; c->R1[2] += c->Qz[1]; }
      double mx = 0, my = 0, mz = 0;
      if (!(e == 0 || e == Z - 2)) { mx = c->Qx[e] + c->Qx[e + 1]; my = c->Qy[e] + c->Qy[e + 1]; mz = c->Qz[e] + c->Qz[e + 1]; }
      const double w_next = c->w[e + 1];
      for (int i = e + 1; i < Z - 1; ++i) mv64(c, i, i + 1);
      c->w[e] = w_next;
      c->Z = Z - 1;
      set64(c, e, Nm, mx, my, mz);
      kind = ORC_EV_CDD;
      break;
    }
    pr -= wd;
    if (pr < wp) { c->N[e] += 1; c->N[e + 1] -= 1; kind = ORC_EV_SHIFT_P; break; }
    pr -= wp;
    if (pr < wm) { c->N[e] -= 1; c->N[e + 1] += 1; kind = ORC_EV_SHIFT_M; break; }
    pr -= wm;
    /* CD creation in strand e */
    NEED_G64();
    const int N = c->N[e];
    const int newn = split_newn_d(pr, N, wc);
    const double wnew = orc_pcd_rate_fdt_d(pcd_fdt(p, x[3]), u_tau);
    for (int i = Z; i > e; --i) mv64(c, i, i - 1);
    c->Z = Z + 1;
    if (e == 0) {
      const double sig = sqrt((double)(N - newn) / 3.0);
      set64(c, 0, newn, 0, 0, 0);
      c->w[0] = wnew;
      set64(c, 1, N - newn, g[0] * sig, g[1] * sig, g[2] * sig);
      if (p->msd) { c->R1[0] -= c->Qx[1]; c->R1[1] -= c->Qy[1]; c->R1[2] -= c->Qz[1]; }
    } else {
      const double sig = sqrt((double)newn * (double)(N - newn) / (3.0 * N));
      const double ratio = (double)newn / N;
      const double qx = c->Qx[e + 1], qy = c->Qy[e + 1], qz = c->Qz[e + 1];
      const double nx = g[0] * sig + qx * ratio, ny = g[1] * sig + qy * ratio, nz = g[2] * sig + qz * ratio;
      set64(c, e, newn, nx, ny, nz);
      c->w[e] = wnew;
      set64(c, e + 1, N - newn, qx - nx, qy - ny, qz - nz);
    }
    kind = ORC_EV_CDC;
  }
  if (kind == ORC_EV_NONE) {
    /* end events in reference order: tail CDc, SDc head, SDc tail, SDd head, SDd tail */
    if (pr < endr[0]) {
      NEED_G64();
      const int N = c->N[Z - 1];
      const int newn = split_newn_d(pr, N, endr[0]);
      const double sig = (Z == 1) ? 0.0 : sqrt((double)newn / 3.0);
      c->Z = Z + 1;
      set64(c, Z - 1, newn, g[0] * sig, g[1] * sig, g[2] * sig);
      c->w[Z - 1] = orc_pcd_rate_fdt_d(pcd_fdt(p, x[3]), u_tau);
      set64(c, Z, N - newn, 0, 0, 0);
      c->w[Z] = 0.0;
      kind = ORC_EV_CDC_TAIL;
    } else if ((pr -= endr[0]) < endr[1]) {
      NEED_G64();
      const int N0 = c->N[0];
      const double sig = (Z == 1) ? 0.0 : sqrt((double)(N0 - 1) / 3.0);
      for (int i = Z; i > 0; --i) mv64(c, i, i - 1);
      c->Z = Z + 1;
      set64(c, 0, 1, 0, 0, 0);
      c->w[0] = p->cd ? orc_pcd_rate_ft_d(pcd_ft(p, x[3]), u_tau) : 0.0;
      set64(c, 1, N0 - 1, g[0] * sig, g[1] * sig, g[2] * sig);
      if (p->msd) { c->R1[0] -= c->Qx[1]; c->R1[1] -= c->Qy[1]; c->R1[2] -= c->Qz[1]; }
      kind = ORC_EV_SDC_HEAD;
    } else if ((pr -= endr[1]) < endr[2]) {
      NEED_G64();
      const int Nz = c->N[Z - 1];
      const double sig = (Z == 1) ? 0.0 : sqrt((double)(Nz - 1) / 3.0);
      c->Z = Z + 1;
      set64(c, Z - 1, Nz - 1, g[0] * sig, g[1] * sig, g[2] * sig);
      c->w[Z - 1] = p->cd ? orc_pcd_rate_ft_d(pcd_ft(p, x[3]), u_tau) : 0.0;
      set64(c, Z, 1, 0, 0, 0);
      c->w[Z] = 0.0;
      kind = ORC_EV_SDC_TAIL;
    } else if ((pr -= endr[2]) < endr[3]) {
      const int N1 = c->N[1];
      if (p->msd) { c->R1[0] += c->Qx[1]; c->R1[1] += c->Qy[1]; c->R1[2] += c->Qz[1]; }
      for (int i = 0; i < Z - 1; ++i) mv64(c, i, i + 1);
      c->Z = Z - 1;
      set64(c, 0, N1 + 1, 0, 0, 0);
      kind = ORC_EV_SDD_HEAD;
    } else if ((pr -= endr[3]) < endr[4]) {
      const int Nz2 = c->N[Z - 2];
      c->Z = Z - 1;
      set64(c, Z - 2, Nz2 + 1, 0, 0, 0);
      c->w[Z - 2] = 0.0;
      kind = ORC_EV_SDD_TAIL;
    }
  }
#undef NEED_G64
  c->ev += 1;
  c->last_event = kind;
  return kind;
}

void orc_chain64_stress_sample(const orc_chain64* c, double out[3]) {
  double a = 0, b = 0, d = 0;
  for (int i = 0; i < c->Z; ++i) {
    const double N = c->N[i];
    a -= 3.0 * c->Qx[i] * c->Qy[i] / N;
    b -= 3.0 * c->Qy[i] * c->Qz[i] / N;
    d -= 3.0 * c->Qx[i] * c->Qz[i] / N;
  }
  out[0] = a; out[1] = b; out[2] = d;
}

void orc_chain64_com_sample(const orc_chain64* c, double out[3]) {
  double P[3] = {0, 0, 0}, s[3] = {0, 0, 0};
  const double* q[3] = {c->Qx, c->Qy, c->Qz};
  for (int i = 0; i < c->Z; ++i)
    for (int k = 0; k < 3; ++k) {
      s[k] += (P[k] + 0.5 * q[k][i]) * c->N[i] / c->nk;
      P[k] += q[k][i];
    }
  for (int k = 0; k < 3; ++k) out[k] = s[k] + c->R1[k];
}

void orc_chain64_stress_full(const orc_chain64* c, double out[8]) {
  double s[7] = {0, 0, 0, 0, 0, 0, 0};
  for (int i = 0; i < c->Z; ++i) {
    const double x = c->Qx[i], y = c->Qy[i], z = c->Qz[i], N = c->N[i];
    s[0] -= 3 * x * x / N; s[1] -= 3 * y * y / N; s[2] -= 3 * z * z / N;
    s[3] -= 3 * x * y / N; s[4] -= 3 * y * z / N; s[5] -= 3 * x * z / N;
    s[6] += sqrt(x * x + y * y + z * z);
  }
  for (int k = 0; k < 7; ++k) out[k] = s[k];
  out[7] = c->Z;
}

static void deform64(orc_chain64* c, const orc_params* p) {
  const double dt = c->dt_prev;
  const double* k = NULL; double kk[9];
  for (int i = 0; i < 9; ++i) kk[i] = p->kappa[i];
  k = kk;
  for (int i = 0; i < c->Z; ++i) {
    const double x = c->Qx[i], y = c->Qy[i], z = c->Qz[i];
    c->Qx[i] = x + dt * (k[0] * x + k[1] * y + k[2] * z);
    c->Qy[i] = y + dt * (k[3] * x + k[4] * y + k[5] * z);
    c->Qz[i] = z + dt * (k[6] * x + k[7] * y + k[8] * z);
  }
}

int64_t orc_chain64_advance(orc_chain64* c, const orc_params* p, const uint32_t key[2], double t_end,
                            int64_t max_events, double delta, int64_t k0, int64_t k_end, double* samples,
                            int64_t stride) {
  int64_t n = 0;
  for (;;) {
    if (p->flow) {
      if (c->dt_prev != 0.0) deform64(c, p);
      c->dt_prev = 0.0;
    }
    if (samples && !p->flow_samples) {
      while (c->next_sample < k_end && c->t > (double)c->next_sample * delta) {
        double sv[3];
        if (p->msd) orc_chain64_com_sample(c, sv);
        else orc_chain64_stress_sample(c, sv);
        if (c->next_sample >= k0) {
          double* o = samples + (c->next_sample - k0) * stride;
          o[0] = sv[0]; o[1] = sv[1]; o[2] = sv[2];
        }
        c->next_sample++;
      }
    } else if (samples) {
      while (c->next_sample < k_end && c->t >= (double)c->next_sample * delta) {
        double sv[8];
        orc_chain64_stress_full(c, sv);
        if (c->next_sample >= k0) {
          double* o = samples + (c->next_sample - k0) * stride;
          for (int q = 0; q < 8; ++q) o[q] = sv[q];
        }
        c->next_sample++;
      }
    }
    if (c->stall || c->t >= t_end || n >= max_events) break;
    orc_chain64_event(c, p, key);
    ++n;
  }
  return n;
}

/* ------------------------------------------------------------------ */
/* multi-tau correlator, FP64 (common/eq_ensemble_kernel.cu:133-227)  */
/* ------------------------------------------------------------------ */
int orc_corr_levels(double sim_time, double delta, int p, int m) {
  /* common/main_cuda.cu:70 */
  return (int)ceil(log(sim_time / delta / p) / log((double)m)) + 1;
}

int orc_corr_init(orc_corr* cr, int p, int m, int s, int nchains) {
  memset(cr, 0, sizeof(*cr));
  cr->p = p; cr->m = m; cr->s = s; cr->nchains = nchains; cr->type = 0;
  cr->ins = calloc(s, sizeof(int)); cr->macc = calloc(s, sizeof(int));
  cr->nins = calloc(s, sizeof(int64_t));
  cr->D = calloc((size_t)nchains * s * p * 3, sizeof(double));
  cr->A = calloc((size_t)nchains * s * 3, sizeof(double));
  cr->C = calloc((size_t)s * p, sizeof(double));
  cr->n = calloc((size_t)s * p, sizeof(double));
  return (cr->D && cr->A && cr->C) ? 0 : -1;
}

void orc_corr_set_type(orc_corr* cr, int type) { cr->type = type; }

void orc_corr_free(orc_corr* cr) {
  free(cr->ins); free(cr->macc); free(cr->nins); free(cr->D); free(cr->A); free(cr->C); free(cr->n);
  memset(cr, 0, sizeof(*cr));
}

/* level-k insert for all chains at once (counters are shared) */
static void corr_add_level(orc_corr* cr, double* w, int k) {
  if (k == cr->s) return;
  const int p = cr->p, m = cr->m, s = cr->s, nc = cr->nchains;
  const int ins = cr->ins[k];
  for (int c = 0; c < nc; ++c) {
    double* D = cr->D + (((size_t)c * s + k) * p) * 3;
    double* A = cr->A + ((size_t)c * s + k) * 3;
    /* type 1 (MSD) carries the first value of each block up, unaveraged
       (common/eq_ensemble_kernel.cu:181-189) */
    for (int q = 0; q < 3; ++q) {
      D[ins * 3 + q] = w[c * 3 + q];
      if (cr->type == 0 || cr->macc[k] == 0) A[q] += w[c * 3 + q];
    }
  }
  cr->macc[k]++;
  if (cr->macc[k] == m) {
    double* up = malloc(sizeof(double) * 3 * (size_t)nc);
    for (int c = 0; c < nc; ++c) {
      double* A = cr->A + ((size_t)c * s + k) * 3;
      for (int q = 0; q < 3; ++q) { up[c * 3 + q] = cr->type == 0 ? A[q] / m : A[q]; A[q] = 0.0; }
    }
    cr->macc[k] = 0;
    corr_add_level(cr, up, k + 1);
    free(up);
  }
  const int jmin = (k == 0) ? 0 : p / m;
  const int64_t have = cr->nins[k];  /* previous inserts at this level */
  for (int j = jmin; j < p; ++j) {
    if ((int64_t)j > have) continue;
    int idx = ins - j;
    if (idx < 0) idx += p;
    double sum = 0.0;
    for (int c = 0; c < nc; ++c) {
      const double* D = cr->D + (((size_t)c * s + k) * p) * 3;
      if (cr->type == 0) {
        sum += D[ins * 3 + 0] * D[idx * 3 + 0] + D[ins * 3 + 1] * D[idx * 3 + 1] + D[ins * 3 + 2] * D[idx * 3 + 2];
      } else {  /* squared displacement (common/eq_ensemble_kernel.cu:203,218) */
        const double ex = D[ins * 3 + 0] - D[idx * 3 + 0], ey = D[ins * 3 + 1] - D[idx * 3 + 1];
        const double ez = D[ins * 3 + 2] - D[idx * 3 + 2];
        sum += ex * ex + ey * ey + ez * ez;
      }
    }
    cr->C[k * p + j] += sum;
    cr->n[k * p + j] += 1.0;
  }
  cr->ins[k] = (ins + 1 == p) ? 0 : ins + 1;
  cr->nins[k]++;
}

void orc_corr_add_row(orc_corr* cr, const double* w) {
  double* tmp = malloc(sizeof(double) * 3 * (size_t)cr->nchains);
  memcpy(tmp, w, sizeof(double) * 3 * (size_t)cr->nchains);
  corr_add_level(cr, tmp, 0);
  cr->nsamples++;
  free(tmp);
}

/* common/correlator.cu:104-187 and common/ensemble.cu:318-321 */
int orc_corr_result(const orc_corr* cr, double* lag, double* g) {
  int im = 0;
  for (int k = 0; k < cr->s; ++k) {
    const int jmin = (k == 0) ? 0 : cr->p / cr->m;
    double mk = 1.0;
    for (int q = 0; q < k; ++q) mk *= cr->m;
    for (int j = jmin; j < cr->p; ++j) {
      const double n = cr->n[k * cr->p + j];
      lag[im] = (double)j * mk;
      /* G: trace/3 (common/correlator.cu:176); MSD: the plain sum of the three components */
      g[im] = n > 0 ? cr->C[k * cr->p + j] / n / (cr->type == 0 ? 3.0 : 1.0) / cr->nchains : 0.0;
      ++im;
    }
  }
  return im;
}

/* ------------------------------------------------------------------ */
/* CPU baseline                                                        */
/* ------------------------------------------------------------------ */
typedef struct {
  const orc_params* p; const uint32_t* key; uint32_t c0, c1; double t_end; int fp64; int64_t events;
} bench_job;

static void* bench_worker(void* arg) {
  bench_job* j = (bench_job*)arg;
  const int cap = j->p->nmax + 2;
  if (j->fp64) {
    orc_chain64 c;
    orc_chain64_alloc(&c, cap);
    for (uint32_t id = j->c0; id < j->c1; ++id) {
      orc_chain64_init_eq(&c, j->p, j->key, id, j->p->nk);
      j->events += orc_chain64_advance(&c, j->p, j->key, j->t_end, INT64_MAX, 1.0, 0, 0, NULL, 0);
    }
    orc_chain64_free(&c);
  } else {
    orc_chain c;
    orc_chain_alloc(&c, cap);
    for (uint32_t id = j->c0; id < j->c1; ++id) {
      orc_chain_init_eq(&c, j->p, j->key, id, j->p->nk);
      j->events += orc_chain_advance(&c, j->p, j->key, j->t_end, INT64_MAX, 1.0, 0, 0, NULL, 0);
    }
    orc_chain_free(&c);
  }
  return NULL;
}

int64_t orc_bench(const orc_params* p, const uint32_t key[2], int n_chains, uint32_t chain0,
                  double t_end, int n_threads, int fp64, double* seconds) {
  if (n_threads < 1) n_threads = 1;
  pthread_t th[256];
  bench_job jobs[256];
  if (n_threads > 256) n_threads = 256;
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  for (int t = 0; t < n_threads; ++t) {
    jobs[t].p = p; jobs[t].key = key; jobs[t].t_end = t_end; jobs[t].fp64 = fp64; jobs[t].events = 0;
    jobs[t].c0 = chain0 + (uint32_t)((int64_t)n_chains * t / n_threads);
    jobs[t].c1 = chain0 + (uint32_t)((int64_t)n_chains * (t + 1) / n_threads);
    pthread_create(&th[t], NULL, bench_worker, &jobs[t]);
  }
  int64_t ev = 0;
  for (int t = 0; t < n_threads; ++t) { pthread_join(th[t], NULL); ev += jobs[t].events; }
  clock_gettime(CLOCK_MONOTONIC, &b);
  *seconds = (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
  return ev;
}

/* ------------------------------------------------------------------ */
/* long-horizon G(t) per chain                                         */
/* ------------------------------------------------------------------ */
typedef struct {
  const orc_params* p; const uint32_t* key; uint32_t chain0; const int* nk; int c0, c1;
  double delta; int64_t nsamples, window; int pm, mm, s, fp64, np; double* out; int64_t events;
} gt_job;

static void* gt_worker(void* arg) {
  gt_job* j = (gt_job*)arg;
  const int cap = j->p->nmax + 2;
  double* buf = malloc(sizeof(double) * 3 * (size_t)j->window);
  float* bf = malloc(sizeof(float) * 3 * (size_t)j->window);
  double* lag = malloc(sizeof(double) * (size_t)j->np);
  orc_chain c32;
  orc_chain64 c64;
  if (j->fp64) orc_chain64_alloc(&c64, cap); else orc_chain_alloc(&c32, cap);
  for (int i = j->c0; i < j->c1; ++i) {
    const uint32_t id = j->chain0 + (uint32_t)i;
    const int nk = j->nk ? j->nk[i] : j->p->nk;
    if (j->fp64) orc_chain64_init_eq(&c64, j->p, j->key, id, nk);
    else orc_chain_init_eq(&c32, j->p, j->key, id, nk);
    orc_corr cr;
    orc_corr_init(&cr, j->pm, j->mm, j->s, 1);
    for (int64_t k0 = 0; k0 < j->nsamples; k0 += j->window) {
      const int64_t k1 = k0 + j->window < j->nsamples ? k0 + j->window : j->nsamples;
      const double te = (double)k1 * j->delta;
      if (j->fp64) {
        j->events += orc_chain64_advance(&c64, j->p, j->key, te, INT64_MAX, j->delta, k0, k1, buf, 3);
      } else {
        j->events += orc_chain_advance(&c32, j->p, j->key, te, INT64_MAX, j->delta, k0, k1, bf, 3);
        for (int64_t q = 0; q < 3 * (k1 - k0); ++q) buf[q] = (double)bf[q];
      }
      for (int64_t k = 0; k < k1 - k0; ++k) orc_corr_add_row(&cr, buf + 3 * k);
    }
    orc_corr_result(&cr, lag, j->out + (size_t)i * j->np);
    orc_corr_free(&cr);
  }
  if (j->fp64) orc_chain64_free(&c64); else orc_chain_free(&c32);
  free(buf); free(bf); free(lag);
  return NULL;
}

int64_t orc_gt_chains(const orc_params* p, const uint32_t key[2], int n, uint32_t chain0, const int* nk_per_chain,
                      double delta, int64_t nsamples, int64_t window, int pm, int mm, int s, int n_threads, int fp64,
                      double* out, double* seconds) {
  if (n_threads < 1) n_threads = 1;
  if (n_threads > 256) n_threads = 256;
  if (n_threads > n) n_threads = n > 0 ? n : 1;
  if (window < 1) window = 256;
  pthread_t th[256];
  gt_job jobs[256];
  struct timespec a, b;
  clock_gettime(CLOCK_MONOTONIC, &a);
  const int np = pm + (s - 1) * (pm - pm / mm);
  for (int t = 0; t < n_threads; ++t) {
    gt_job* j = &jobs[t];
    j->p = p; j->key = key; j->chain0 = chain0; j->nk = nk_per_chain; j->delta = delta;
    j->nsamples = nsamples; j->window = window; j->pm = pm; j->mm = mm; j->s = s; j->fp64 = fp64; j->np = np;
    j->out = out; j->events = 0;
    j->c0 = (int)((int64_t)n * t / n_threads);
    j->c1 = (int)((int64_t)n * (t + 1) / n_threads);
    pthread_create(&th[t], NULL, gt_worker, j);
  }
  int64_t ev = 0;
  for (int t = 0; t < n_threads; ++t) { pthread_join(th[t], NULL); ev += jobs[t].events; }
  clock_gettime(CLOCK_MONOTONIC, &b);
  *seconds = (double)(b.tv_sec - a.tv_sec) + 1e-9 * (double)(b.tv_nsec - a.tv_nsec);
  return ev;
}

/* ------------------------------------------------------------------ */
/* ensemble API                                                        */
/* ------------------------------------------------------------------ */
struct orc_ens {
  const orc_params* p;
  uint32_t key[2];
  int n, cap, fp64;
  orc_chain* c32;
  orc_chain64* c64;
};

orc_params* orc_params_new(float beta, int nk, int nmax, int cd, const float* kappa, int exp_wait) {
  orc_params* p = malloc(sizeof(orc_params));
  if (orc_params_init(p, beta, nk, nmax, cd, kappa, exp_wait)) { free(p); return NULL; }
  return p;
}
void orc_params_delete(orc_params* p) { if (p) { orc_params_free(p); free(p); } }

void orc_params_set_flow_samples(orc_params* p, int on) { p->flow_samples = on ? 1 : 0; }
void orc_params_set_msd(orc_params* p, int on) { p->msd = on ? 1 : 0; }

int orc_params_table(const orc_params* p, int which, float* out) {
  const float* t = NULL;
  int n = p->nmax + 2;
  switch (which) {
    case 0: t = p->SP; break; case 1: t = p->SM; break; case 2: t = p->PP; break;
    case 3: t = p->PM; break; case 4: t = p->EC; break; case 5: t = p->ED; break;
    case 6: t = p->ED2; break; case 7: t = p->T3; break; case 8: t = p->CC; break;
    case 9: t = p->FR; n = p->fr_size; break;
    case 10: case 11: {
      if (!p->partner) return 0;
      const orc_pcd* q = which == 10 ? p->pf : p->pdd;
      for (int k = 0; k < ORC_PARTNER_K && out; ++k) {
        float* o = out + 5 * k;
        if (which == 10) { o[0] = q[k].s_At; o[1] = q[k].s_Ct; o[2] = q[k].s_Dt; o[3] = q[k].s_taud_inv; o[4] = q[k].s_one_minus_g; }
        else { o[0] = q[k].s_Adt; o[1] = q[k].s_Bdt; o[2] = q[k].s_Cdt; o[3] = q[k].s_Ddt; o[4] = q[k].s_taud_inv; }
      }
      return 5 * ORC_PARTNER_K;
    }
    case 12: if (!p->partner) return 0; t = p->pf_nk; n = ORC_PARTNER_K; break;
    case 13: if (!p->partner) return 0; t = p->pd_nk; n = ORC_PARTNER_K; break;
    default: return 0;
  }
  if (out) memcpy(out, t, sizeof(float) * (size_t)n);
  return n;
}

void orc_params_scalars(const orc_params* p, float* o) {
  const orc_pcd* q = &p->pcd;
  const float v[24] = {q->g, q->alpha, q->tau_0, q->tau_max, q->tau_d, q->c1, q->cm1, q->rcm1c1,
                       q->At, q->Adt, q->Bdt, q->normdt, q->s_At, q->s_Ct, q->s_Dt, q->s_Adt,
                       q->s_Bdt, q->s_Cdt, q->s_Ddt, q->s_taud_inv, q->s_one_minus_g, p->c_create,
                       (float)p->cut_threshold, p->beta};
  memcpy(o, v, sizeof(v));
}

orc_ens* orc_ens_new(const orc_params* p, uint32_t key0, uint32_t key1, uint32_t chain0, int nchains,
                     const int* nk_per_chain, int fp64) {
  orc_ens* e = calloc(1, sizeof(orc_ens));
  e->p = p; e->key[0] = key0; e->key[1] = key1; e->n = nchains; e->cap = p->nmax + 2; e->fp64 = fp64;
  if (fp64) {
    e->c64 = calloc((size_t)nchains, sizeof(orc_chain64));
    for (int i = 0; i < nchains; ++i) {
      orc_chain64_alloc(&e->c64[i], e->cap);
      orc_chain64_init_eq(&e->c64[i], p, e->key, chain0 + (uint32_t)i, nk_per_chain ? nk_per_chain[i] : p->nk);
    }
  } else {
    e->c32 = calloc((size_t)nchains, sizeof(orc_chain));
    for (int i = 0; i < nchains; ++i) {
      orc_chain_alloc(&e->c32[i], e->cap);
      orc_chain_init_eq(&e->c32[i], p, e->key, chain0 + (uint32_t)i, nk_per_chain ? nk_per_chain[i] : p->nk);
    }
  }
  return e;
}

void orc_ens_delete(orc_ens* e) {
  if (!e) return;
  for (int i = 0; i < e->n; ++i) {
    if (e->fp64) orc_chain64_free(&e->c64[i]); else orc_chain_free(&e->c32[i]);
  }
  free(e->c32); free(e->c64); free(e);
}

typedef struct {
  orc_ens* e; int c0, c1; double t_end; int64_t max_events; double delta; int64_t k0, k_end;
  void* samples; int64_t events; int ncomp;
} ens_job;

static void* ens_worker(void* arg) {
  ens_job* j = (ens_job*)arg;
  orc_ens* e = j->e;
  const int64_t stride = (int64_t)e->n * j->ncomp;
  for (int i = j->c0; i < j->c1; ++i) {
    if (e->fp64) {
      double* s = j->samples ? (double*)j->samples + (int64_t)i * j->ncomp : NULL;
      j->events += orc_chain64_advance(&e->c64[i], e->p, e->key, j->t_end, j->max_events, j->delta, j->k0, j->k_end, s, stride);
    } else {
      float* s = j->samples ? (float*)j->samples + (int64_t)i * j->ncomp : NULL;
      j->events += orc_chain_advance(&e->c32[i], e->p, e->key, j->t_end, j->max_events, j->delta, j->k0, j->k_end, s, stride);
    }
  }
  return NULL;
}

int64_t orc_ens_advance(orc_ens* e, double t_end, int64_t max_events, double delta, int64_t k0,
                        int64_t k_end, void* samples, int n_threads) {
  if (n_threads < 1) n_threads = 1;
  if (n_threads > 64) n_threads = 64;
  if (n_threads > e->n) n_threads = e->n > 0 ? e->n : 1;
  pthread_t th[64];
  ens_job jobs[64];
  for (int t = 0; t < n_threads; ++t) {
    jobs[t].e = e; jobs[t].t_end = t_end; jobs[t].max_events = max_events; jobs[t].delta = delta;
    jobs[t].k0 = k0; jobs[t].k_end = k_end; jobs[t].samples = samples; jobs[t].events = 0;
    jobs[t].ncomp = e->p->flow_samples ? 8 : 3;
    jobs[t].c0 = (int)((int64_t)e->n * t / n_threads);
    jobs[t].c1 = (int)((int64_t)e->n * (t + 1) / n_threads);
    pthread_create(&th[t], NULL, ens_worker, &jobs[t]);
  }
  int64_t ev = 0;
  for (int t = 0; t < n_threads; ++t) { pthread_join(th[t], NULL); ev += jobs[t].events; }
  return ev;
}

int orc_ens_cap(const orc_ens* e) { return e->cap; }

void orc_ens_get_r1(const orc_ens* e, float* r1) {
  for (int i = 0; i < e->n; ++i)
    for (int k = 0; k < 3; ++k) r1[i * 3 + k] = e->fp64 ? (float)e->c64[i].R1[k] : e->c32[i].R1[k];
}

void orc_ens_get(const orc_ens* e, int* Z, int* N, float* Q, float* w, double* t, uint64_t* ev, int* stall) {
  const int cap = e->cap;
  for (int i = 0; i < e->n; ++i) {
    if (e->fp64) {
      const orc_chain64* c = &e->c64[i];
      if (Z) Z[i] = c->Z;
      if (t) t[i] = c->t;
      if (ev) ev[i] = c->ev;
      if (stall) stall[i] = c->stall;
      for (int k = 0; k < cap; ++k) {
        const int on = k < c->Z;
        if (N) N[(size_t)i * cap + k] = on ? c->N[k] : 0;
        if (Q) {
          Q[((size_t)i * cap + k) * 3 + 0] = on ? (float)c->Qx[k] : 0.0f;
          Q[((size_t)i * cap + k) * 3 + 1] = on ? (float)c->Qy[k] : 0.0f;
          Q[((size_t)i * cap + k) * 3 + 2] = on ? (float)c->Qz[k] : 0.0f;
        }
        if (w) w[(size_t)i * cap + k] = (k < c->Z - 1) ? (float)c->w[k] : 0.0f;
      }
    } else {
      const orc_chain* c = &e->c32[i];
      if (Z) Z[i] = c->Z;
      if (t) t[i] = c->t;
      if (ev) ev[i] = c->ev;
      if (stall) stall[i] = c->stall;
      for (int k = 0; k < cap; ++k) {
        const int on = k < c->Z;
        if (N) N[(size_t)i * cap + k] = on ? c->N[k] : 0;
        if (Q) {
          Q[((size_t)i * cap + k) * 3 + 0] = on ? c->Qx[k] : 0.0f;
          Q[((size_t)i * cap + k) * 3 + 1] = on ? c->Qy[k] : 0.0f;
          Q[((size_t)i * cap + k) * 3 + 2] = on ? c->Qz[k] : 0.0f;
        }
        if (w) w[(size_t)i * cap + k] = (k < c->Z - 1) ? c->w[k] : 0.0f;
      }
    }
  }
}

void orc_ens_get64(const orc_ens* e, double* Q, double* w) {
  if (!e->fp64) return;
  const int cap = e->cap;
  for (int i = 0; i < e->n; ++i) {
    const orc_chain64* c = &e->c64[i];
    for (int k = 0; k < cap; ++k) {
      const int on = k < c->Z;
      Q[((size_t)i * cap + k) * 3 + 0] = on ? c->Qx[k] : 0.0;
      Q[((size_t)i * cap + k) * 3 + 1] = on ? c->Qy[k] : 0.0;
      Q[((size_t)i * cap + k) * 3 + 2] = on ? c->Qz[k] : 0.0;
      w[(size_t)i * cap + k] = (k < c->Z - 1) ? c->w[k] : 0.0;
    }
  }
}

void orc_ens_set(orc_ens* e, const int* Z, const int* N, const float* Q, const float* w, const double* t,
                 const uint64_t* ev, const float* dt_prev, const int64_t* next_sample) {
  if (e->fp64) return;
  const int cap = e->cap;
  for (int i = 0; i < e->n; ++i) {
    orc_chain* c = &e->c32[i];
    c->Z = Z[i];
    for (int k = 0; k < cap; ++k) {
      c->N[k] = N[(size_t)i * cap + k];
      c->Qx[k] = Q[((size_t)i * cap + k) * 3 + 0];
      c->Qy[k] = Q[((size_t)i * cap + k) * 3 + 1];
      c->Qz[k] = Q[((size_t)i * cap + k) * 3 + 2];
      c->w[k] = w[(size_t)i * cap + k];
    }
    if (t) c->t = t[i];
    if (ev) c->ev = ev[i];
    if (dt_prev) c->dt_prev = dt_prev[i];
    if (next_sample) c->next_sample = next_sample[i];
    c->stall = 0;
    orc_chain_refresh(c, e->p);
  }
}

void orc_ens_stress(const orc_ens* e, double* out) {
  for (int i = 0; i < e->n; ++i) {
    if (e->fp64) {
      orc_chain64_stress_full(&e->c64[i], out + (size_t)i * 8);
    } else {
      float v[8];
      orc_chain_stress_full(&e->c32[i], e->p, v);
      for (int k = 0; k < 8; ++k) out[(size_t)i * 8 + k] = v[k];
    }
  }
}

/* bulk evaluation of the mirror FP32 math: which 0 exp, 1 log, 2 sin(2 pi v), 3 cos(2 pi v) */
void orc_vec_math(int which, const float* in, float* out, int n) {
  for (int i = 0; i < n; ++i) {
    float s, c;
    switch (which) {
      case 0: out[i] = orc_expf(in[i]); break;
      case 1: out[i] = orc_logf(in[i]); break;
      case 2: orc_sincos2pif(in[i], &s, &c); out[i] = s; break;
      default: orc_sincos2pif(in[i], &s, &c); out[i] = c; break;
    }
  }
}

/* standalone pair rates for checks: FP32 table form (mirror) and the reference
   formula in double (common/eq_ensemble_kernel.cu:97-124) */
void orc_pair_rate_f32(const orc_params* p, int Ni, int Nj, float q2i, float q2j, int endi, int endj,
                       float* wp, float* wm) {
  orc_chain c;
  memset(&c, 0, sizeof(c));
  int N[3] = {Ni, Nj, 1};
  float Q2[3] = {q2i, q2j, 0.0f}, zero[3] = {0, 0, 0}, Wp[3], Wm[3], R[3];
  c.N = N; c.Q2 = Q2; c.w = zero; c.Wp = Wp; c.Wm = Wm; c.R = R;
  /* end flags are positional: pair 0 of a 2-strand chain has both ends */
  if (endi && endj) { c.Z = 2; pair_rate(&c, p, 0); *wp = Wp[0]; *wm = Wm[0]; return; }
  int N4[4], k = 0;
  float Q4[4];
  float z4[4] = {0, 0, 0, 0}, Wp4[4], Wm4[4], R4[4];
  if (!endi) { N4[k] = 1; Q4[k] = 0.0f; ++k; }
  const int e = k;
  N4[k] = Ni; Q4[k] = q2i; ++k;
  N4[k] = Nj; Q4[k] = q2j; ++k;
  if (!endj) { N4[k] = 1; Q4[k] = 0.0f; ++k; }
  c.Z = k; c.N = N4; c.Q2 = Q4; c.w = z4; c.Wp = Wp4; c.Wm = Wm4; c.R = R4;
  pair_rate(&c, p, e);
  *wp = Wp4[e]; *wm = Wm4[e];
}

/* largest |cached - recomputed| over every cached pair rate (checks the dirty-window logic) */
double orc_ens_cache_error(const orc_ens* e) {
  double worst = 0.0;
  if (e->fp64) return 0.0;
  for (int i = 0; i < e->n; ++i) {
    orc_chain* c = &e->c32[i];
    for (int k = 0; k < c->Z - 1; ++k) {
      const float wp = c->Wp[k], wm = c->Wm[k], r = c->R[k];
      const float q2 = c->Q2[k];
      q2_of(c, k);
      if (c->Q2[k] != q2) worst = worst > 1e30 ? worst : 1e30;
      pair_rate(c, e->p, k);
      const double d = fabs((double)wp - c->Wp[k]) + fabs((double)wm - c->Wm[k]) + fabs((double)r - c->R[k]);
      if (d > worst) worst = d;
    }
  }
  return worst;
}

/* invariant check of every chain: returns the number of violations
   (sum N != nk, N < 1, end Q != 0, Z outside 1..nk) */
int orc_ens_invariants(const orc_ens* e) {
  int bad = 0;
  for (int i = 0; i < e->n; ++i) {
    if (e->fp64) {
      const orc_chain64* c = &e->c64[i];
      long s = 0;
      if (c->Z < 1 || c->Z > c->nk) ++bad;
      for (int k = 0; k < c->Z; ++k) { s += c->N[k]; if (c->N[k] < 1) ++bad; }
      if (s != c->nk) ++bad;
      if (c->Qx[0] != 0 || c->Qy[0] != 0 || c->Qz[0] != 0) ++bad;
      if (c->Qx[c->Z - 1] != 0 || c->Qy[c->Z - 1] != 0 || c->Qz[c->Z - 1] != 0) ++bad;
    } else {
      const orc_chain* c = &e->c32[i];
      long s = 0;
      if (c->Z < 1 || c->Z > c->nk) ++bad;
      for (int k = 0; k < c->Z; ++k) { s += c->N[k]; if (c->N[k] < 1) ++bad; }
      if (s != c->nk) ++bad;
      if (c->Qx[0] != 0 || c->Qy[0] != 0 || c->Qz[0] != 0) ++bad;
      if (c->Qx[c->Z - 1] != 0 || c->Qy[c->Z - 1] != 0 || c->Qz[c->Z - 1] != 0) ++bad;
    }
  }
  return bad;
}

float orc_params_rate_ft(const orc_params* p, float u) { return orc_pcd_rate_ft(&p->pcd, u); }
float orc_params_rate_fdt(const orc_params* p, float u) { return orc_pcd_rate_fdt(&p->pcd, u); }
