"""ctypes binding over the C ABI in include/dsm.h (libgpu_dsm.so, built for gfx950).

The shared library is the product; this module only marshals arguments.  It
raises immediately when the library is missing or cannot run on the device --
there is no CPU fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libgpu_dsm.so")
EXE_PATH = os.path.join(HERE, "_build", "gpu_DSM")

SAMPLE = {"none": 0, "gt": 1, "flow": 2, "msd": 3}
MODE = {"batched": 0, "plain": 1}
TABLES = ["SP", "SM", "PP", "PM", "EC", "ED", "ED2", "T3", "CC", "FR", "PF", "PD", "PF_NK", "PD_NK"]
SCALARS = ["g", "alpha", "tau_0", "tau_max", "tau_d", "c1", "cm1", "rcm1c1", "At", "Adt", "Bdt",
           "normdt", "s_At", "s_Ct", "s_Dt", "s_Adt", "s_Bdt", "s_Cdt", "s_Ddt", "s_taud_inv",
           "s_one_minus_g", "c_create", "cut_threshold", "beta"]
EXPORTS = ["dsm_create", "dsm_destroy", "dsm_last_error", "dsm_mem_info", "dsm_add_ensemble",
           "dsm_remove_ensemble", "dsm_init_equilibrium", "dsm_state_cap", "dsm_get_state", "dsm_set_state", "dsm_get_r1",
           "dsm_advance", "dsm_advance_many", "dsm_advance_events", "dsm_flow_stress", "dsm_correlator",
           "dsm_last_samples", "dsm_tables", "dsm_ensemble_scalars", "dsm_probe_scan", "dsm_probe_math",
           "dsm_probe_philox", "dsm_probe_com", "dsm_probe_copy", "dsm_main", "dsm_main_ex", "dsm_serve", "dsm_serve_plan",
           "dsm_ensemble_info", "dsm_save", "dsm_load", "dsm_load_error", "dsm_stress", "dsm_check_state",
           "dsm_serve_ex", "dsm_correlator_range", "dsm_set_option", "dsm_get_option", "dsm_stress_noalloc"]
OPTIONS = {"plain_jumps": 1, "small_per_cu": 2, "split_margin": 3, "lds_table_floats": 4, "concurrent": 5,
           "blocks_per_cu": 6, "window": 7, "graph": 8, "wide_pair": 9}
CHECKS = ["z_range", "n_below_1", "n_sum", "end_q", "nonfinite", "stalled", "behind"]


class DsmError(RuntimeError):
    pass


class EnsembleDesc(C.Structure):
    _fields_ = [("n_chains", C.c_int32), ("nk", C.c_int32), ("nk_per_chain", C.POINTER(C.c_int32)),
                ("beta", C.c_float), ("cd", C.c_int32), ("kappa", C.c_float * 9), ("exp_wait", C.c_int32),
                ("seed", C.c_uint32), ("salt", C.c_uint32), ("sample_mode", C.c_int32), ("delta", C.c_double),
                ("corr_p", C.c_int32), ("corr_m", C.c_int32), ("corr_levels", C.c_int32), ("window", C.c_int32),
                ("partner_mode", C.c_int32), ("gex", C.c_float * 4)]

PARTNER = {None: 0, "own": 0, "gex": 1, "melt": 2}


class RunStats(C.Structure):
    _fields_ = [("events", C.c_uint64), ("z_sum", C.c_uint64), ("kmc_ms", C.c_double),
                ("first_kmc_ms", C.c_double), ("total_ms", C.c_double), ("kmc_launches", C.c_int64),
                ("first_launches", C.c_int64)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


class ServeConfig(C.Structure):
    _fields_ = [("n_requests", C.c_int32), ("arrival_rate", C.c_double), ("seed", C.c_uint64),
                ("nk_min", C.c_int32), ("nk_max", C.c_int32), ("nk_step", C.c_int32), ("n_beta", C.c_int32),
                ("beta", C.c_float * 4), ("nch_min", C.c_int32), ("nch_max", C.c_int32),
                ("flow_fraction", C.c_double), ("rate_min", C.c_double), ("rate_max", C.c_double),
                ("horizon", C.c_double), ("window", C.c_double), ("cd", C.c_int32), ("hbm_fraction", C.c_double),
                ("progress_s", C.c_double), ("graph", C.c_int32)]


class ServeReport(C.Structure):
    _fields_ = [("completed", C.c_int32), ("max_active", C.c_int32), ("rounds", C.c_int32),
                ("makespan_s", C.c_double), ("p50_s", C.c_double), ("p99_s", C.c_double), ("mean_s", C.c_double),
                ("max_s", C.c_double), ("jumps", C.c_uint64), ("jumps_per_s", C.c_double),
                ("peak_bytes", C.c_int64), ("admit_s", C.c_double), ("advance_s", C.c_double),
                ("fetch_s", C.c_double)]

    def as_dict(self):
        return {f: getattr(self, f) for f, _ in self._fields_}


# default serving stream: 2000 requests at 20/s, Nk in 16..512, beta in {0.5, 1, 2},
# 1024..65536 chains, half of them sheared at 1e-5..1e-2 per tau_K, 1e5 tau_K each
# (about 1.7e14 jumps in total; callers pass a shorter horizon for timed runs)
SERVE_DEFAULTS = dict(n_requests=2000, arrival_rate=20.0, seed=1, nk_min=16, nk_max=512, nk_step=16,
                      beta=(0.5, 1.0, 2.0), nch_min=1024, nch_max=65536, flow_fraction=0.5, rate_min=1e-5,
                      rate_max=1e-2, horizon=1.0e5, window=10.0, cd=1, hbm_fraction=0.8,
                      progress_s=0.0, graph=0)


class DeviceChains(C.Structure):
    """dsm_device_chains: device addresses of a chain state in the caller's memory."""
    _fields_ = [(f, C.c_void_p) for f in ("Z", "stall", "base", "dt_prev", "Qx", "Qy", "Qz", "N")]


class MainOptions(C.Structure):
    _fields_ = [("mode", C.c_int32), ("corr_p", C.c_int32), ("corr_m", C.c_int32), ("window", C.c_int32),
                ("exp_wait", C.c_int32), ("quiet", C.c_int32), ("workdir", C.c_char_p)]


_lib = None


def load_library():
    """Load libgpu_dsm.so; raises DsmError if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise DsmError("gpu_dsm_amd HIP library not built (%s); run __graft_entry__.build()" % LIB_PATH)
    L = C.CDLL(LIB_PATH)
    P = C.POINTER
    L.dsm_last_error.restype = C.c_char_p
    L.dsm_create.argtypes = [C.c_int32, P(C.c_void_p)]
    L.dsm_destroy.argtypes = [C.c_void_p]
    L.dsm_mem_info.argtypes = [C.c_int32, P(C.c_size_t), P(C.c_size_t)]
    L.dsm_add_ensemble.argtypes = [C.c_void_p, P(EnsembleDesc), P(C.c_int32)]
    L.dsm_remove_ensemble.argtypes = [C.c_void_p, C.c_int32]
    L.dsm_init_equilibrium.argtypes = [C.c_void_p, C.c_int32]
    L.dsm_state_cap.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32)]
    L.dsm_get_state.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_float),
                                P(C.c_float), P(C.c_double), P(C.c_uint64), P(C.c_int32), P(C.c_float),
                                P(C.c_int64)]
    L.dsm_set_state.argtypes = L.dsm_get_state.argtypes
    L.dsm_get_r1.argtypes = [C.c_void_p, C.c_int32, P(C.c_float)]
    L.dsm_ensemble_info.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32), P(C.c_int32), P(C.c_int32)]
    L.dsm_save.argtypes = [C.c_void_p, C.c_int32, C.c_char_p, C.c_double]
    L.dsm_load.argtypes = [C.c_void_p, C.c_int32, C.c_char_p]
    L.dsm_load_error.restype = C.c_char_p
    L.dsm_stress.argtypes = [C.c_void_p, C.c_int32, P(C.c_double), P(C.c_int32)]
    L.dsm_check_state.argtypes = [C.c_void_p, C.c_int32, C.c_double, P(C.c_int64)]
    L.dsm_stress_noalloc.argtypes = [C.c_int32, P(DeviceChains), C.c_int32, P(C.c_float), P(C.c_double),
                                     P(C.c_int32)]
    L.dsm_set_option.argtypes = [C.c_void_p, C.c_int32, C.c_int32]
    L.dsm_get_option.argtypes = [C.c_void_p, C.c_int32, P(C.c_int32)]
    L.dsm_serve_ex.argtypes = [C.c_int32, P(ServeConfig), P(ServeReport), P(C.c_double), P(C.c_double)]
    L.dsm_advance.argtypes = [C.c_void_p, C.c_int32, C.c_double, C.c_int32, P(C.c_int32), P(C.c_int32),
                              P(RunStats)]
    L.dsm_advance_many.argtypes = [C.c_void_p, P(C.c_int32), P(C.c_double), C.c_int32, C.c_int32, P(C.c_int32),
                                   P(RunStats)]
    L.dsm_advance_events.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.c_int32, P(RunStats)]
    L.dsm_flow_stress.argtypes = [C.c_void_p, C.c_int32, C.c_int64, C.c_int64, P(C.c_double), P(C.c_int64)]
    L.dsm_correlator.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_double), P(C.c_double), P(C.c_int32)]
    L.dsm_correlator_range.argtypes = [C.c_void_p, C.c_int32, C.c_int32, C.c_int32, C.c_int32, P(C.c_double),
                                       P(C.c_double), P(C.c_int32)]
    L.dsm_last_samples.argtypes = [C.c_void_p, C.c_int32, P(C.c_float), C.c_int64, P(C.c_int64), P(C.c_int64)]
    L.dsm_tables.argtypes = [C.c_void_p, C.c_int32, C.c_int32, P(C.c_float), P(C.c_int32)]
    L.dsm_ensemble_scalars.argtypes = [C.c_void_p, C.c_int32, P(C.c_float)]
    L.dsm_probe_scan.argtypes = [C.c_int32, P(C.c_float), P(C.c_float), C.c_int32, C.c_int32]
    L.dsm_probe_math.argtypes = [C.c_int32, P(C.c_float), P(C.c_float), C.c_int32, C.c_int32]
    L.dsm_probe_philox.argtypes = [C.c_int32, P(C.c_uint32), P(C.c_uint32), C.c_int32]
    L.dsm_probe_com.argtypes = [C.c_int32, P(C.c_float), P(C.c_float), C.c_int32, C.c_int32]
    L.dsm_probe_copy.argtypes = [C.c_int32, C.c_int64, C.c_int32, P(C.c_double)]
    L.dsm_serve.argtypes = [C.c_int32, P(ServeConfig), P(ServeReport), P(C.c_double)]
    L.dsm_serve_plan.argtypes = [P(ServeConfig), C.c_int32, P(C.c_double), P(C.c_int32), P(C.c_float),
                                 P(C.c_int32), P(C.c_double), P(C.c_int64)]
    L.dsm_main_ex.argtypes = [P(C.c_bool), C.c_int, C.c_char_p, C.c_char_p, C.c_int, C.c_bool, P(C.c_int),
                              P(MainOptions)]
    _lib = L
    return L


def _check(rc):
    if rc != 0:
        msg = load_library().dsm_last_error().decode(errors="replace")
        raise DsmError("gpu_dsm_amd error %d: %s" % (rc, msg))


def _p(a, ct):
    return a.ctypes.data_as(C.POINTER(ct))


def mem_info(device=0):
    f, t = C.c_size_t(), C.c_size_t()
    _check(load_library().dsm_mem_info(device, C.byref(f), C.byref(t)))
    return f.value, t.value


class Engine:
    """One GPU engine holding any number of ensembles (handle over dsm_engine)."""

    def __init__(self, device=0):
        self._L = load_library()
        h = C.c_void_p()
        _check(self._L.dsm_create(int(device), C.byref(h)))
        self.h = h
        self.device = device
        self.ens = {}

    def close(self):
        if getattr(self, "h", None):
            self._L.dsm_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    # ------------------------------------------------------------ options
    def set_option(self, name, value):
        _check(self._L.dsm_set_option(self.h, OPTIONS[name], int(value)))

    def get_option(self, name):
        v = C.c_int32()
        _check(self._L.dsm_get_option(self.h, OPTIONS[name], C.byref(v)))
        return v.value

    def options(self):
        return {k: self.get_option(k) for k in OPTIONS}

    # ------------------------------------------------------------ ensembles
    def add_ensemble(self, n_chains, nk, beta=1.0, cd=0, kappa=None, seed=1, salt=0, sample="none", delta=1.0,
                     corr_p=64, corr_m=8, corr_levels=0, window=0, exp_wait=False, nk_per_chain=None,
                     partner=None, gex=None):
        """partner: None/"own" (CD partners have the chain's Nk), "gex" (GEX weight
        distribution with gex = (a, b, mp, Mk), the reference's PD_flag 1) or "melt" (the
        ensemble's own chain lengths, weight-averaged)."""
        d = EnsembleDesc()
        d.n_chains, d.nk, d.beta, d.cd = int(n_chains), int(nk), float(beta), int(cd)
        kap = np.zeros(9, np.float32) if kappa is None else np.asarray(kappa, np.float32)
        for i in range(9):
            d.kappa[i] = float(kap[i])
        d.exp_wait = int(bool(exp_wait))
        d.seed, d.salt = int(seed) & 0xFFFFFFFF, int(salt) & 0xFFFFFFFF
        d.sample_mode = SAMPLE[sample]
        d.delta = float(delta)
        d.corr_p, d.corr_m, d.corr_levels, d.window = int(corr_p), int(corr_m), int(corr_levels), int(window)
        d.partner_mode = PARTNER[partner]
        if gex is not None:
            for i in range(4):
                d.gex[i] = float(gex[i])
        keep = None
        if nk_per_chain is not None:
            keep = np.ascontiguousarray(nk_per_chain, np.int32)
            d.nk_per_chain = _p(keep, C.c_int32)
        eid = C.c_int32()
        _check(self._L.dsm_add_ensemble(self.h, C.byref(d), C.byref(eid)))
        cap = C.c_int32()
        _check(self._L.dsm_state_cap(self.h, eid.value, C.byref(cap)))
        self.ens[eid.value] = dict(n=int(n_chains), cap=cap.value, sample=sample)
        return eid.value

    def remove_ensemble(self, ens):
        _check(self._L.dsm_remove_ensemble(self.h, int(ens)))
        self.ens.pop(ens, None)

    def init_equilibrium(self, ens):
        _check(self._L.dsm_init_equilibrium(self.h, int(ens)))

    def state(self, ens, cap=None):
        n = self.ens[ens]["n"]
        cap = cap or self.ens[ens]["cap"]
        Z = np.zeros(n, np.int32); N = np.zeros((n, cap), np.int32)
        Q = np.zeros((n, cap, 3), np.float32); w = np.zeros((n, cap), np.float32)
        t = np.zeros(n); ev = np.zeros(n, np.uint64); st = np.zeros(n, np.int32)
        dt = np.zeros(n, np.float32); ns = np.zeros(n, np.int64)
        _check(self._L.dsm_get_state(self.h, int(ens), int(cap), _p(Z, C.c_int32), _p(N, C.c_int32),
                                     _p(Q, C.c_float), _p(w, C.c_float), _p(t, C.c_double), _p(ev, C.c_uint64),
                                     _p(st, C.c_int32), _p(dt, C.c_float), _p(ns, C.c_int64)))
        out = dict(Z=Z, N=N, Q=Q, w=w, t=t, ev=ev, stall=st, dt_prev=dt, next_sample=ns)
        if self.ens[ens]["sample"] == "msd":
            r1 = np.zeros((n, 3), np.float32)
            _check(self._L.dsm_get_r1(self.h, int(ens), _p(r1, C.c_float)))
            out["R1"] = r1
        return out

    def set_state(self, ens, Z, N, Q, w, t=None, ev=None, stall=None, dt_prev=None, next_sample=None):
        n = self.ens[ens]["n"]
        cap = N.shape[1]
        arrs = [np.ascontiguousarray(Z, np.int32), np.ascontiguousarray(N, np.int32),
                np.ascontiguousarray(Q, np.float32), np.ascontiguousarray(w, np.float32)]
        opt = [(t, np.float64, C.c_double), (ev, np.uint64, C.c_uint64), (stall, np.int32, C.c_int32),
               (dt_prev, np.float32, C.c_float), (next_sample, np.int64, C.c_int64)]
        ptrs = [_p(arrs[0], C.c_int32), _p(arrs[1], C.c_int32), _p(arrs[2], C.c_float), _p(arrs[3], C.c_float)]
        keep = []
        for a, dt, ct in opt:
            if a is None:
                ptrs.append(None)
            else:
                x = np.ascontiguousarray(a, dt)
                keep.append(x)
                ptrs.append(_p(x, ct))
        assert arrs[0].shape[0] == n
        _check(self._L.dsm_set_state(self.h, int(ens), int(cap), *ptrs))

    def stress(self, ens):
        """Ensemble stress of the current state: (xx yy zz xy yz zx Lpp Z averages, healthy chains)."""
        out = np.zeros(8)
        h = C.c_int32()
        _check(self._L.dsm_stress(self.h, int(ens), _p(out, C.c_double), C.byref(h)))
        return out, h.value

    def check(self, ens, t_min=-1.0):
        """Invariant violations of the ensemble's state counted on the device (dict by CHECKS)."""
        out = np.zeros(7, np.int64)
        _check(self._L.dsm_check_state(self.h, int(ens), float(t_min), _p(out, C.c_int64)))
        return dict(zip(CHECKS, out.tolist()))

    def save(self, ens, path, t_ref=0.0):
        """Write the -s binary file of the ensemble (times relative to t_ref)."""
        _check(self._L.dsm_save(self.h, int(ens), str(path).encode(), float(t_ref)))

    def load(self, ens, path):
        """Read a -s / -l binary file (reference layout) into the ensemble."""
        rc = self._L.dsm_load(self.h, int(ens), str(path).encode())
        if rc != 0:
            msg = self._L.dsm_load_error().decode(errors="replace") or self._L.dsm_last_error().decode(errors="replace")
            raise DsmError("gpu_dsm_amd error %d: %s" % (rc, msg))

    def advance(self, ens, t_end, mode="batched", run_flag=None):
        rs = RunStats()
        flag = None
        if run_flag is not None:
            flag = C.c_int32(int(run_flag))
        rc = self._L.dsm_advance(self.h, int(ens), float(t_end), MODE[mode],
                                 C.byref(flag) if flag is not None else None, None, C.byref(rs))
        if rc == -1:
            return dict(rs.as_dict(), aborted=True)
        _check(rc)
        return rs.as_dict()

    def advance_many(self, ens_ids, t_end, mode="batched"):
        ids = np.ascontiguousarray(ens_ids, np.int32)
        te = np.ascontiguousarray(np.broadcast_to(np.asarray(t_end, np.float64), ids.shape), np.float64)
        rs = RunStats()
        _check(self._L.dsm_advance_many(self.h, _p(ids, C.c_int32), _p(te, C.c_double), len(ids), MODE[mode],
                                        None, C.byref(rs)))
        return rs.as_dict()

    def advance_events(self, ens, n_events, mode="batched"):
        rs = RunStats()
        _check(self._L.dsm_advance_events(self.h, int(ens), int(n_events), MODE[mode], C.byref(rs)))
        return rs.as_dict()

    def flow_stress(self, ens):
        n = C.c_int64()
        _check(self._L.dsm_flow_stress(self.h, int(ens), 0, 0, None, C.byref(n)))
        out = np.zeros((n.value, 9))
        if n.value:
            _check(self._L.dsm_flow_stress(self.h, int(ens), 0, n.value, _p(out, C.c_double), C.byref(n)))
        return out

    def correlator(self, ens):
        npts = C.c_int32()
        _check(self._L.dsm_correlator(self.h, int(ens), 0, None, None, C.byref(npts)))
        lag = np.zeros(npts.value); g = np.zeros(npts.value)
        _check(self._L.dsm_correlator(self.h, int(ens), npts.value, _p(lag, C.c_double), _p(g, C.c_double),
                                      C.byref(npts)))
        return lag, g

    def correlator_chains(self, ens, first=0, count=None, group=1):
        """G(t) of consecutive chain groups: returns (lags, G[n_groups][np]), group g
        holding chains first + g*group .. first + (g+1)*group - 1."""
        n = self.ens[ens]["n"]
        count = n - first if count is None else count
        npts = C.c_int32()
        _check(self._L.dsm_correlator(self.h, int(ens), 0, None, None, C.byref(npts)))
        ng = count // group
        lag = np.zeros(npts.value)
        out = np.zeros((ng, npts.value))
        for k in range(ng):
            g = np.zeros(npts.value)
            _check(self._L.dsm_correlator_range(self.h, int(ens), int(first + k * group), int(group), npts.value,
                                                _p(lag, C.c_double), _p(g, C.c_double), C.byref(npts)))
            out[k] = g
        return lag, out

    def last_samples(self, ens):
        k0, rows = C.c_int64(), C.c_int64()
        _check(self._L.dsm_last_samples(self.h, int(ens), None, 0, C.byref(k0), C.byref(rows)))
        n = self.ens[ens]["n"]
        comp = 8 if self.ens[ens]["sample"] == "flow" else 3
        out = np.zeros((rows.value, n, comp), np.float32)
        if rows.value:
            _check(self._L.dsm_last_samples(self.h, int(ens), _p(out, C.c_float), out.size, C.byref(k0),
                                            C.byref(rows)))
        return k0.value, out

    def table(self, ens, name):
        n = C.c_int32()
        which = TABLES.index(name)
        _check(self._L.dsm_tables(self.h, int(ens), which, None, C.byref(n)))
        out = np.zeros(n.value, np.float32)
        _check(self._L.dsm_tables(self.h, int(ens), which, _p(out, C.c_float), C.byref(n)))
        return out

    def scalars(self, ens):
        out = np.zeros(24, np.float32)
        _check(self._L.dsm_ensemble_scalars(self.h, int(ens), _p(out, C.c_float)))
        return dict(zip(SCALARS, out.tolist()))


# ---------------------------------------------------------------- building-block probes
def probe_com(Q, N, geom, device=0):
    """MSD centre-of-mass sample (relative to R1) of one chain state computed by the
    KMC kernels' code in geometry `geom` (0 row form, 1 pair form, 2/3/4 wave form
    G = 1/2/4, 7 pair form with two register groups, 8 row form with two register groups)."""
    Q = np.asarray(Q, np.float32).reshape(-1, 3)
    N = np.asarray(N)
    n = Q.shape[0]
    inp = np.zeros(4 * n + 1, np.float32)
    inp[:4 * n] = np.concatenate([Q, N[:, None].astype(np.float32)], axis=1).ravel()
    inp[4 * n] = float(N.sum())
    out = np.zeros(3, np.float32)
    _check(load_library().dsm_probe_com(device, _p(inp, C.c_float), _p(out, C.c_float), n, int(geom)))
    return out


def probe_scan(values, form=1, device=0):
    """Canonical scan tree on the device.  form 1: one chain (len <= 1024) -> (inclusive
    sums, total).  form 0: values shaped (4, n<=64), four row-form chains -> (sums (4, n),
    totals (4,)); form 2: values shaped (2, n<=128), two pair-form chains; form 3: (2, n<=256),
    two chains of the pair form with two register groups; form 4: (4, n<=128), four chains of
    the row form with two register groups; form 5: as form 1 (len <= 512) in the wave form
    with two register groups."""
    v = np.ascontiguousarray(values, np.float32)
    if form in (1, 5):
        n = v.size
        out = np.zeros(n + 1, np.float32)
        _check(load_library().dsm_probe_scan(device, _p(v, C.c_float), _p(out, C.c_float), n, int(form)))
        return out[:n], out[n]
    rows = 2 if form in (2, 3) else 4  # forms 0 and 4: four chains
    v = v.reshape(rows, -1)
    n = v.shape[1]
    out = np.zeros((rows, n + 1), np.float32)
    _check(load_library().dsm_probe_scan(device, _p(v, C.c_float), _p(out, C.c_float), n, int(form)))
    return out[:, :n], out[:, n]


def probe_math(values, which, device=0):
    v = np.ascontiguousarray(values, np.float32)
    out = np.zeros_like(v)
    _check(load_library().dsm_probe_math(device, _p(v, C.c_float), _p(out, C.c_float), len(v),
                                         {"exp": 0, "log": 1, "sin2pi": 2, "cos2pi": 3}[which]))
    return out


def probe_philox(ctr_key, device=0):
    ck = np.ascontiguousarray(ctr_key, np.uint32).reshape(-1, 6)
    out = np.zeros((ck.shape[0], 4), np.uint32)
    _check(load_library().dsm_probe_philox(device, _p(ck, C.c_uint32), _p(out, C.c_uint32), ck.shape[0]))
    return out


def stress_noalloc(addresses, n_chains, kappa=None, device=0):
    """dsm_stress_noalloc over a chain state in device memory the caller owns:
    `addresses` maps Z, stall, base, dt_prev, Qx, Qy, Qz, N to device addresses (ints).
    Returns (xx yy zz xy yz zx Lpp Z averages, healthy chains)."""
    dc = DeviceChains(*[int(addresses[f]) for f, _ in DeviceChains._fields_])
    k = np.zeros(9, np.float32) if kappa is None else np.ascontiguousarray(kappa, np.float32)
    out = np.zeros(8)
    h = C.c_int32()
    _check(load_library().dsm_stress_noalloc(int(device), C.byref(dc), int(n_chains), _p(k, C.c_float),
                                             _p(out, C.c_double), C.byref(h)))
    return out, h.value


def probe_copy(n_bytes, reps=20, device=0):
    """Milliseconds per plain device copy of n_bytes (read n_bytes + write n_bytes)."""
    ms = C.c_double(0.0)
    _check(load_library().dsm_probe_copy(device, int(n_bytes), int(reps), C.byref(ms)))
    return ms.value


def _serve_config(**kw):
    c = dict(SERVE_DEFAULTS)
    c.update(kw)
    cfg = ServeConfig()
    for k, v in c.items():
        if k == "beta":
            b = list(v)
            cfg.n_beta = len(b)
            for i, x in enumerate(b):
                cfg.beta[i] = float(x)
        else:
            setattr(cfg, k, v)
    return cfg, c


def serve_plan(**kw):
    """The request stream of a serving run (host only): dict of per-request arrays."""
    cfg, c = _serve_config(**kw)
    n = c["n_requests"]
    out = dict(arrival_s=np.zeros(n), nk=np.zeros(n, np.int32), beta=np.zeros(n, np.float32),
               n_chains=np.zeros(n, np.int32), shear_rate=np.zeros(n), bytes=np.zeros(n, np.int64))
    _check(load_library().dsm_serve_plan(C.byref(cfg), n, _p(out["arrival_s"], C.c_double), _p(out["nk"], C.c_int32),
                                         _p(out["beta"], C.c_float), _p(out["n_chains"], C.c_int32),
                                         _p(out["shear_rate"], C.c_double), _p(out["bytes"], C.c_int64)))
    return out


def serve(device=0, summary=False, **kw):
    """Run a serving stream on the device; returns (report dict, per-request latency in s)
    and, with summary=True, each request's result summary (DSM_SERVE_SUMMARY = 18 doubles:
    kind 1 G(t) / 2 flow, points or rows, then G at lags 0..15 or the last flow row)."""
    cfg, c = _serve_config(**kw)
    rep = ServeReport()
    lat = np.zeros(c["n_requests"])
    sm = np.zeros((c["n_requests"], 18)) if summary else None
    _check(load_library().dsm_serve_ex(device, C.byref(cfg), C.byref(rep), _p(lat, C.c_double),
                                       _p(sm, C.c_double) if summary else None))
    if summary:
        return rep.as_dict(), lat, sm
    return rep.as_dict(), lat


def run_main(job_id=0, savefile=None, loadfile=None, device=0, distr=False, mode="batched", corr_p=64, corr_m=8,
             window=0, exp_wait=False, quiet=True, workdir=None):
    """Drop-in driver (dsm_main_ex): same behaviour as gpu_DSM with these arguments."""
    L = load_library()
    flag = C.c_bool(True)
    prog = C.c_int(0)
    o = MainOptions(MODE[mode], int(corr_p), int(corr_m), int(window), int(bool(exp_wait)), int(bool(quiet)),
                    workdir.encode() if workdir else None)
    rc = L.dsm_main_ex(C.byref(flag), int(job_id), savefile.encode() if savefile else None,
                       loadfile.encode() if loadfile else None, int(device), bool(distr), C.byref(prog),
                       C.byref(o))
    return rc
