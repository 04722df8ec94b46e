"""ctypes wrapper around the CPU oracle (oracle/dsm_oracle.c).

Test infrastructure only: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker.  The product package
(gpu_dsm_amd) never imports this module.
"""
import ctypes as C
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libdsm_oracle.so")

PURPOSE_STEP, PURPOSE_GAUSS = 0, 1
PURPOSE_INIT_CUT, PURPOSE_INIT_Q, PURPOSE_INIT_TAU, PURPOSE_INIT_NK = 16, 17, 18, 19
EVENT_NAMES = ["none", "shift+", "shift-", "CDd", "CDc", "CDc_tail", "SDc_head", "SDc_tail",
               "SDd_head", "SDd_tail", "stall"]
TABLES = ["SP", "SM", "PP", "PM", "EC", "ED", "ED2", "T3", "CC", "FR", "PF", "PD", "PF_NK", "PD_NK"]
SCALARS = ["g", "alpha", "tau_0", "tau_max", "tau_d", "c1", "cm1", "rcm1c1", "At", "Adt", "Bdt",
           "normdt", "s_At", "s_Ct", "s_Dt", "s_Adt", "s_Bdt", "s_Cdt", "s_Ddt", "s_taud_inv",
           "s_one_minus_g", "c_create", "cut_threshold", "beta"]

_lib = None


def build():
    """Compile the oracle with its Makefile (gcc)."""
    import subprocess
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = C.CDLL(_LIB_PATH)
        f32p, f64p = C.POINTER(C.c_float), C.POINTER(C.c_double)
        u32 = C.c_uint32
        L.orc_philox4x32_10.argtypes = [C.POINTER(u32), C.POINTER(u32), C.POINTER(u32)]
        L.orc_u01.argtypes = [u32]; L.orc_u01.restype = C.c_float
        for n in ("orc_expf", "orc_logf"):
            getattr(L, n).argtypes = [C.c_float]; getattr(L, n).restype = C.c_float
        L.orc_powf.argtypes = [C.c_float, C.c_float]; L.orc_powf.restype = C.c_float
        L.orc_sincos2pif.argtypes = [C.c_float, f32p, f32p]
        L.orc_boxmuller.argtypes = [u32, u32, f32p, f32p]
        L.orc_scan_tree.argtypes = [f32p, C.c_int, f32p]
        L.orc_tree_total.argtypes = [f32p, C.c_int]
        L.orc_tree_total.restype = C.c_float
        L.orc_vec_math.argtypes = [C.c_int, f32p, f32p, C.c_int]
        L.orc_params_new.argtypes = [C.c_float, C.c_int, C.c_int, C.c_int, f32p, C.c_int]
        L.orc_params_new.restype = C.c_void_p
        L.orc_params_delete.argtypes = [C.c_void_p]
        L.orc_params_set_flow_samples.argtypes = [C.c_void_p, C.c_int]
        L.orc_params_set_msd.argtypes = [C.c_void_p, C.c_int]
        L.orc_params_set_partner.argtypes = [C.c_void_p, C.c_int, f32p, C.POINTER(C.c_int), C.c_int]
        L.orc_gamma_p.argtypes = [C.c_double, C.c_double]; L.orc_gamma_p.restype = C.c_double
        L.orc_ens_get_r1.argtypes = [C.c_void_p, f32p]
        L.orc_params_table.argtypes = [C.c_void_p, C.c_int, f32p]; L.orc_params_table.restype = C.c_int
        L.orc_params_scalars.argtypes = [C.c_void_p, f32p]
        L.orc_params_rate_ft.argtypes = [C.c_void_p, C.c_float]; L.orc_params_rate_ft.restype = C.c_float
        L.orc_params_rate_fdt.argtypes = [C.c_void_p, C.c_float]; L.orc_params_rate_fdt.restype = C.c_float
        L.orc_ens_new.argtypes = [C.c_void_p, u32, u32, u32, C.c_int, C.POINTER(C.c_int), C.c_int]
        L.orc_ens_new.restype = C.c_void_p
        L.orc_ens_delete.argtypes = [C.c_void_p]
        L.orc_ens_advance.argtypes = [C.c_void_p, C.c_double, C.c_int64, C.c_double, C.c_int64,
                                      C.c_int64, C.c_void_p, C.c_int]
        L.orc_ens_advance.restype = C.c_int64
        L.orc_ens_cap.argtypes = [C.c_void_p]; L.orc_ens_cap.restype = C.c_int
        L.orc_ens_get.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), f32p, f32p, f64p,
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_int)]
        L.orc_ens_get64.argtypes = [C.c_void_p, f64p, f64p]
        L.orc_ens_set.argtypes = [C.c_void_p, C.POINTER(C.c_int), C.POINTER(C.c_int), f32p, f32p, f64p,
                                  C.POINTER(C.c_uint64), f32p, C.POINTER(C.c_int64)]
        L.orc_ens_stress.argtypes = [C.c_void_p, f64p]
        L.orc_pair_rate_f32.argtypes = [C.c_void_p, C.c_int, C.c_int, C.c_float, C.c_float, C.c_int, C.c_int,
                                        f32p, f32p]
        L.orc_ens_cache_error.argtypes = [C.c_void_p]; L.orc_ens_cache_error.restype = C.c_double
        L.orc_ens_invariants.argtypes = [C.c_void_p]; L.orc_ens_invariants.restype = C.c_int
        L.orc_corr_levels.argtypes = [C.c_double, C.c_double, C.c_int, C.c_int]
        L.orc_corr_levels.restype = C.c_int
        L.orc_bench.argtypes = [C.c_void_p, C.POINTER(u32), C.c_int, u32, C.c_double, C.c_int, C.c_int, f64p]
        L.orc_bench.restype = C.c_int64
        L.orc_gt_chains.argtypes = [C.c_void_p, C.POINTER(u32), C.c_int, u32, C.POINTER(C.c_int), C.c_double,
                                    C.c_int64, C.c_int64, C.c_int, C.c_int, C.c_int, C.c_int, C.c_int, f64p, f64p]
        L.orc_gt_chains.restype = C.c_int64
        _lib = L
    return _lib


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*[int(x) & 0xFFFFFFFF for x in ctr])
    k = (C.c_uint32 * 2)(*[int(x) & 0xFFFFFFFF for x in key])
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return [int(v) for v in o]


def expf(x):
    return lib().orc_expf(float(x))


def logf(x):
    return lib().orc_logf(float(x))


def powf(x, y):
    return lib().orc_powf(float(x), float(y))


def sincos2pi(v):
    s, c = C.c_float(), C.c_float()
    lib().orc_sincos2pif(float(v), C.byref(s), C.byref(c))
    return s.value, c.value


def u01(x):
    return lib().orc_u01(int(x) & 0xFFFFFFFF)


def boxmuller(a, b):
    g0, g1 = C.c_float(), C.c_float()
    lib().orc_boxmuller(int(a) & 0xFFFFFFFF, int(b) & 0xFFFFFFFF, C.byref(g0), C.byref(g1))
    return g0.value, g1.value


def vec_math(which, values):
    """Mirror FP32 exp / log / sin(2 pi v) / cos(2 pi v) over an array."""
    v = np.ascontiguousarray(values, np.float32)
    out = np.zeros_like(v)
    lib().orc_vec_math({"exp": 0, "log": 1, "sin2pi": 2, "cos2pi": 3}[which], _p(v, C.c_float),
                       _p(out, C.c_float), len(v))
    return out


def scan_tree(v):
    v = np.ascontiguousarray(v, dtype=np.float32)
    out = np.zeros_like(v)
    lib().orc_scan_tree(_p(v, C.c_float), len(v), _p(out, C.c_float))
    return out


def tree_total(v):
    """Total of the canonical scan tree (G at the end of the last occupied 64-block)."""
    v = np.ascontiguousarray(v, dtype=np.float32)
    return lib().orc_tree_total(_p(v, C.c_float), len(v))


def corr_levels(sim_time, delta, p, m):
    return lib().orc_corr_levels(float(sim_time), float(delta), int(p), int(m))


class Params:
    """Ensemble parameters and FP32 rate tables (orc_params)."""

    def __init__(self, beta, nk, cd, kappa=None, nmax=None, exp_wait=False, partner=None, gex=None,
                 nk_chains=None):
        """partner: None (CD partners have the ensemble's Nk), "gex" (GEX weight distribution,
        gex = (a, b, mp, Mk)) or "melt" (weight distribution of nk_chains)."""
        self.beta, self.nk, self.cd = float(beta), int(nk), int(cd)
        self.kappa = np.zeros(9, np.float32) if kappa is None else np.asarray(kappa, np.float32)
        self.nmax = int(nmax if nmax is not None else nk)
        self.h = lib().orc_params_new(self.beta, self.nk, self.nmax, self.cd, _p(self.kappa, C.c_float),
                                      int(bool(exp_wait)))
        if not self.h:
            raise RuntimeError("orc_params_new failed")
        if partner is not None:
            g = np.zeros(4, np.float32) if gex is None else np.asarray(gex, np.float32)
            nkc = np.ascontiguousarray(nk_chains if nk_chains is not None else [self.nk], np.int32)
            mode = {"gex": 1, "melt": 2}[partner]
            if lib().orc_params_set_partner(self.h, mode, _p(g, C.c_float), _p(nkc, C.c_int), nkc.size):
                raise ValueError("bad partner parameters")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_params_delete(self.h)
            self.h = None

    def table(self, name):
        which = TABLES.index(name)
        n = lib().orc_params_table(self.h, which, None)
        out = np.zeros(n, np.float32)
        lib().orc_params_table(self.h, which, _p(out, C.c_float))
        return out

    def scalars(self):
        out = np.zeros(24, np.float32)
        lib().orc_params_scalars(self.h, _p(out, C.c_float))
        return dict(zip(SCALARS, out.tolist()))

    def pair_rates(self, Ni, Nj, q2i, q2j, endi, endj):
        """FP32 (mirror) W+ and W- of a strand pair."""
        wp, wm = C.c_float(), C.c_float()
        lib().orc_pair_rate_f32(self.h, int(Ni), int(Nj), float(q2i), float(q2j), int(endi), int(endj),
                                C.byref(wp), C.byref(wm))
        return wp.value, wm.value

    def rate_ft(self, u):
        return lib().orc_params_rate_ft(self.h, float(u))

    def rate_fdt(self, u):
        return lib().orc_params_rate_fdt(self.h, float(u))


class Ensemble:
    """n chains (ids chain0..chain0+n-1) in mirror-FP32 (fp64=False) or FP64 mode."""

    def __init__(self, params, seed, n, chain0=0, salt=0, fp64=False, nk_per_chain=None):
        self.params, self.n, self.fp64 = params, int(n), bool(fp64)
        self.seed, self.salt, self.chain0 = int(seed), int(salt), int(chain0)
        nkp = None
        if nk_per_chain is not None:
            self._nk = np.ascontiguousarray(nk_per_chain, np.int32)
            nkp = _p(self._nk, C.c_int)
        self.h = lib().orc_ens_new(params.h, self.seed & 0xFFFFFFFF, self.salt & 0xFFFFFFFF,
                                   self.chain0, self.n, nkp, int(self.fp64))
        self.cap = lib().orc_ens_cap(self.h)

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.orc_ens_delete(self.h)
            self.h = None

    def advance(self, t_end=float("inf"), max_events=2**62, delta=1.0, k0=0, k_end=0, samples=False,
                threads=None, flow_samples=False, msd=False):
        """Advance every chain; returns (events, samples[k][chain][ncomp] or None).

        ncomp is 3 (G(t) stress samples, or with msd=True the centre of mass) or 8
        (flow_samples: xx yy zz xy yz zx Lpp Z).  msd=True also tracks R1, the first
        entanglement's fixed-frame position, through the jumps."""
        threads = threads or min(os.cpu_count() or 1, 16)
        lib().orc_params_set_flow_samples(self.params.h, int(bool(flow_samples)))
        lib().orc_params_set_msd(self.params.h, int(bool(msd)))
        buf = None
        if samples and k_end > k0:
            buf = np.zeros((k_end - k0, self.n, 8 if flow_samples else 3), np.float64 if self.fp64 else np.float32)
        ev = lib().orc_ens_advance(self.h, float(t_end), int(max_events), float(delta), int(k0), int(k_end),
                                   buf.ctypes.data_as(C.c_void_p) if buf is not None else None, int(threads))
        return ev, buf

    def state(self):
        n, cap = self.n, self.cap
        Z = np.zeros(n, np.int32); N = np.zeros((n, cap), np.int32)
        Q = np.zeros((n, cap, 3), np.float32); w = np.zeros((n, cap), np.float32)
        t = np.zeros(n, np.float64); ev = np.zeros(n, np.uint64); st = np.zeros(n, np.int32)
        lib().orc_ens_get(self.h, _p(Z, C.c_int), _p(N, C.c_int), _p(Q, C.c_float), _p(w, C.c_float),
                          _p(t, C.c_double), _p(ev, C.c_uint64), _p(st, C.c_int))
        r1 = np.zeros((n, 3), np.float32)
        lib().orc_ens_get_r1(self.h, _p(r1, C.c_float))
        out = dict(Z=Z, N=N, Q=Q, w=w, t=t, ev=ev, stall=st, R1=r1)
        if self.fp64:
            Q64 = np.zeros((n, cap, 3)); w64 = np.zeros((n, cap))
            lib().orc_ens_get64(self.h, _p(Q64, C.c_double), _p(w64, C.c_double))
            out["Q64"], out["w64"] = Q64, w64
        return out

    def set_next_sample(self, k):
        st = self.state()
        self.set_state(st["Z"], st["N"], st["Q"], st["w"], t=st["t"], ev=st["ev"],
                       next_sample=np.full(self.n, k, np.int64))

    def set_state(self, Z, N, Q, w, t=None, ev=None, dt_prev=None, next_sample=None):
        n, cap = self.n, self.cap
        Z = np.ascontiguousarray(Z, np.int32)
        Nf = np.zeros((n, cap), np.int32); Nf[:, :N.shape[1]] = N[:, :cap]
        Qf = np.zeros((n, cap, 3), np.float32); Qf[:, :Q.shape[1]] = Q[:, :cap]
        wf = np.zeros((n, cap), np.float32); wf[:, :w.shape[1]] = w[:, :cap]
        args = [_p(Z, C.c_int), _p(Nf, C.c_int), _p(Qf, C.c_float), _p(wf, C.c_float)]
        keep = []
        for a, dt, ct in ((t, np.float64, C.c_double), (ev, np.uint64, C.c_uint64),
                          (dt_prev, np.float32, C.c_float), (next_sample, np.int64, C.c_int64)):
            if a is None:
                args.append(None)
            else:
                arr = np.ascontiguousarray(a, dt); keep.append(arr); args.append(_p(arr, ct))
        lib().orc_ens_set(self.h, *args)

    def cache_error(self):
        return lib().orc_ens_cache_error(self.h)

    def invariant_violations(self):
        return lib().orc_ens_invariants(self.h)

    def chain_stress(self):
        out = np.zeros((self.n, 8))
        lib().orc_ens_stress(self.h, _p(out, C.c_double))
        return out


def bench(params, seed, n_chains, t_end, threads, fp64=False, chain0=0, salt=0):
    """CPU baseline: returns (events, seconds) for n_chains chains run to t_end."""
    key = (C.c_uint32 * 2)(seed & 0xFFFFFFFF, salt & 0xFFFFFFFF)
    sec = C.c_double()
    ev = lib().orc_bench(params.h, key, int(n_chains), int(chain0), float(t_end), int(threads), int(fp64),
                         C.byref(sec))
    return int(ev), sec.value


def gt_chains(params, seed, n, T, delta=1.0, p=16, m=2, s=None, threads=None, fp64=True, chain0=0, salt=0,
              nk_per_chain=None, window=256):
    """Per-chain G(t) over a long horizon (no stored samples): returns (lags, G[n][np],
    jumps, seconds).  The ensemble G is G.mean(axis=0); chains are independent, so
    error bars follow from the spread over chains."""
    nsamples = int(math.ceil(T / delta - 1e-12))
    if s is None:
        s = corr_levels(T, delta, p, m)
    np_ = p + (s - 1) * (p - p // m)
    out = np.zeros((n, np_))
    key = (C.c_uint32 * 2)(seed & 0xFFFFFFFF, salt & 0xFFFFFFFF)
    sec = C.c_double()
    nkp = None
    if nk_per_chain is not None:
        nkarr = np.ascontiguousarray(nk_per_chain, np.int32)
        nkp = _p(nkarr, C.c_int)
    threads = threads or min(os.cpu_count() or 1, 16)
    ev = lib().orc_gt_chains(params.h, key, int(n), int(chain0), nkp, float(delta), nsamples, int(window), int(p),
                             int(m), int(s), int(threads), int(bool(fp64)), _p(out, C.c_double), C.byref(sec))
    lags = []
    for k in range(s):
        for j in range(0 if k == 0 else p // m, p):
            lags.append(j * m ** k)
    return np.array(lags, np.float64), out, int(ev), sec.value


class Correlator:
    """Multi-tau correlator in FP64 with shared (synchronous) counters, pure numpy.

    Follows common/eq_ensemble_kernel.cu:133-227 (insert, block average,
    lag products) and common/correlator.cu:104-187 (result extraction).
    type 0: stress autocorrelation (block averages go up a level, products, trace/3);
    type 1: MSD (the first value of each block goes up unaveraged, squared
    differences, sum of the three components).
    """

    def __init__(self, p, m, s, nchains, type=0):
        self.p, self.m, self.s, self.nc, self.type = p, m, s, nchains, int(type)
        self.D = np.zeros((s, p, nchains, 3))
        self.A = np.zeros((s, nchains, 3))
        self.ins = np.zeros(s, np.int64)
        self.macc = np.zeros(s, np.int64)
        self.nins = np.zeros(s, np.int64)
        self.C = np.zeros((s, p))
        self.n = np.zeros((s, p))

    def _add(self, w, k):
        if k == self.s:
            return
        p, m = self.p, self.m
        ins = self.ins[k]
        self.D[k, ins] = w
        if self.type == 0 or self.macc[k] == 0:
            self.A[k] += w
        self.macc[k] += 1
        if self.macc[k] == m:
            up = self.A[k] / m if self.type == 0 else self.A[k].copy()
            self.A[k] = 0.0
            self.macc[k] = 0
            self._add(up, k + 1)
        jmin = 0 if k == 0 else p // m
        have = self.nins[k]
        for j in range(jmin, p):
            if j > have:
                continue
            idx = (ins - j) % p
            if self.type == 0:
                self.C[k, j] += float(np.sum(self.D[k, ins] * self.D[k, idx]))
            else:
                d = self.D[k, ins] - self.D[k, idx]
                self.C[k, j] += float(np.sum(d * d))
            self.n[k, j] += 1.0
        self.ins[k] = (ins + 1) % p
        self.nins[k] += 1

    def add_row(self, w):
        self._add(np.asarray(w, np.float64).reshape(self.nc, 3), 0)

    def result(self):
        lags, g = [], []
        for k in range(self.s):
            jmin = 0 if k == 0 else self.p // self.m
            for j in range(jmin, self.p):
                n = self.n[k, j]
                lags.append(j * self.m ** k)
                g.append(self.C[k, j] / n / (3.0 if self.type == 0 else 1.0) / self.nc if n > 0 else 0.0)
        return np.array(lags, np.float64), np.array(g)
