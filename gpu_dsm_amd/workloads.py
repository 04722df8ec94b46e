"""Reference workloads of the DSM ensemble path, one JSON line each.

    python -m gpu_dsm_amd.workloads c1 [--horizon T]      1000 chains, Nk=25, CD off, G(t)
    python -m gpu_dsm_amd.workloads c3 [--horizon T]      262,144 chains, log-normal Nk, CD on, G(t)
    python -m gpu_dsm_amd.workloads c4 [--horizon T]      16 sheared ensembles x 16384 chains, Nk=50
    python -m gpu_dsm_amd.workloads c5 [--horizon T]      serving stream, 2000 requests at 20/s
    python -m gpu_dsm_amd.workloads eta0 [--horizon T]    Green-Kubo viscosity of the C4 chains at rest

bench.py at the repo root measures the headline (the c3 melt, 16 tau_K windows) and the
131,072-chain Nk=100 workload as its secondary line.  Every run starts from the equilibrium draw; all inputs are synthetic
(Philox-keyed); times are wall clock around the device work.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

from . import distributions as dist
from . import engine as eng


def c1(horizon=1.0e5, seed=1, window=256, graph=0):
    """1000 chains, Nk=25, beta=1, CD off, G(t) with p=16, m=2 (the small CLI case).
    `window`: samples (tau_K) per staging window, the host synchronisation interval;
    `graph`: 1 runs the windows as captured HIP graphs."""
    n, nk, p, m = 1000, 25, 16, 2
    s = int(math.ceil(math.log(horizon / p) / math.log(m))) + 1
    with eng.Engine(0) as e:
        e.set_option("graph", graph)
        ens = e.add_ensemble(n, nk, 1.0, 0, seed=seed, sample="gt", delta=1.0, corr_p=p, corr_m=m, corr_levels=s,
                             window=window)
        e.init_equilibrium(ens)
        z0 = e.state(ens)["Z"]
        t0 = time.perf_counter()
        st = e.advance(ens, horizon)
        wall = time.perf_counter() - t0
        lag, g = e.correlator(ens)
    return dict(workload="c1: 1000 chains, Nk=25, beta=1, CD off, G(t) p=16 m=2", horizon_tau_k=horizon,
                window=window, graph=graph,
                jumps=st["events"], seconds=wall, jumps_per_s=st["events"] / wall,
                G0=float(g[0]), G0_from_initial_Z=float(np.mean(np.maximum(z0 - 2, 0))),
                npoints=int(len(g)))


def c3_lengths(n=262144, seed=1, nk_lo=10, nk_hi=2000):
    """Chain lengths of the polydisperse melt: Nk = round(LogNormal(mu, sigma)) with
    number-average 200 and Mw/Mn = 1.5 (sigma^2 = ln 1.5, mu = ln 200 - sigma^2/2),
    clipped to [nk_lo, nk_hi]."""
    s2 = math.log(1.5)
    rng = np.random.default_rng(seed)
    nk = np.rint(rng.lognormal(math.log(200.0) - s2 / 2, math.sqrt(s2), n))
    return np.clip(nk, nk_lo, nk_hi).astype(np.int32)


def c3(horizon=100.0, seed=1, n=262144, distr=None):
    """Polydisperse melt: 262,144 chains of log-normal length (Mn 200, PDI 1.5), beta=1,
    CD on with constraint-release partners drawn from the melt's own length distribution
    (weight-averaged), G(t) with p=16, m=2, all chains in one batched pass per window.
    The output carries the Z / N / Q distributions of the final state against the
    equilibrium laws; `distr` (a directory) also receives them as the reference's
    distr_Z.dat, distr_N.dat and distr_Q.dat tables (the |Q| CDF at 100,000 evenly spaced
    rows instead of one row per internal strand: 2.6e7 here)."""
    nk = c3_lengths(n, seed)
    p, m = 16, 2
    s = int(math.ceil(math.log(horizon / p) / math.log(m))) + 1
    with eng.Engine(0) as e:
        ens = e.add_ensemble(n, int(nk.max()), 1.0, 1, seed=seed, sample="gt", delta=1.0, corr_p=p, corr_m=m,
                             corr_levels=s, nk_per_chain=nk, partner="melt")
        e.init_equilibrium(ens)
        z0 = e.state(ens)["Z"]
        t0 = time.perf_counter()
        st = e.advance(ens, horizon)
        wall = time.perf_counter() - t0
        lag, g = e.correlator(ens)
        Z1, N1, Q1 = dist.fetch_zn_q(e, ens)
    z1 = Z1
    dsum = dist.summarize(Z1, N1, Q1, nk, 1.0)
    if distr:
        os.makedirs(distr, exist_ok=True)
        tables = dist.reference_files(Z1, N1, Q1, max_q_rows=100000)
        for name, (x, y) in zip(("distr_Z.dat", "distr_N.dat", "distr_Q.dat"), tables):
            with open(os.path.join(distr, name), "w") as f:
                f.writelines("%s\t%.9g\n" % (a, b) for a, b in zip(x.tolist(), y.tolist()))
    del N1, Q1
    return dict(workload="c3: %d chains, Nk log-normal Mn 200 PDI 1.5 clipped to [10, %d], beta=1, CD on "
                         "(melt partners), G(t)"
                % (n, int(nk.max())), horizon_tau_k=horizon, jumps=st["events"], seconds=wall,
                jumps_per_s=st["events"] / wall, kmc_ms=st["kmc_ms"], first_kmc_ms=st["first_kmc_ms"],
                mean_Z_start=float(z0.mean()), mean_Z_end=float(z1.mean()), max_Z=int(z1.max()),
                chains_Z_over_64=int((z1 > 64).sum()), G0=float(g[0]), distributions=dsum)


def c4(horizon=1.0e5, seed=1, segment=5000.0):
    """16 ensembles x 16384 chains, Nk=50, beta=1, CD on, shear rates 1e-5..1e-1 per
    tau_K (log-spaced), all advanced in one batched pass per output window.  Runs in
    segments of `segment` tau_K (a progress line on stderr after each).  The steady-state
    stress is the mean over the second half of the run with a standard error from ten
    consecutive blocks; a rate counts as steady when the discarded first half holds at
    least 10 strain units (rate * horizon / 2 >= 10)."""
    n, nk, k = 16384, 50, 16
    rates = [10.0 ** (-5.0 + 4.0 * q / (k - 1)) for q in range(k)]
    with eng.Engine(0) as e:
        ids = []
        for q, gd in enumerate(rates):
            ens = e.add_ensemble(n, nk, 1.0, 1, kappa=[0, gd, 0, 0, 0, 0, 0, 0, 0], seed=seed, salt=q,
                                 sample="flow", delta=1.0, window=64)
            e.init_equilibrium(ens)
            ids.append(ens)
        t0 = time.perf_counter()
        jumps, t = 0, 0.0
        while t < horizon:
            t = min(horizon, t + segment)
            st = e.advance_many(ids, t)
            jumps += st["events"]
            print("c4: t = %g tau_K, %.3g jumps, %.1f s" % (t, jumps, time.perf_counter() - t0), file=sys.stderr,
                  flush=True)
        wall = time.perf_counter() - t0
        out = []
        for ens, gd in zip(ids, rates):
            rows = e.flow_stress(ens)
            half = rows[rows.shape[0] // 2:]
            sxy_t = half[:, 3] / half[:, 8]
            n1_t = (half[:, 0] - half[:, 1]) / half[:, 8]
            blocks = np.array_split(np.arange(half.shape[0]), 10)
            bs = np.array([sxy_t[b].mean() for b in blocks])
            bn = np.array([n1_t[b].mean() for b in blocks])
            out.append(dict(rate=gd, sigma_xy=float(sxy_t.mean()), sigma_xy_se=float(bs.std(ddof=1) / np.sqrt(10)),
                            N1=float(n1_t.mean()), N1_se=float(bn.std(ddof=1) / np.sqrt(10)),
                            viscosity=float(-sxy_t.mean() / gd), steady=bool(gd * horizon / 2 >= 10.0)))
    return dict(workload="c4: 16 x 16384 chains, Nk=50, beta=1, CD on, shear 1e-5..1e-1", horizon_tau_k=horizon,
                jumps=jumps, seconds=wall, jumps_per_s=jumps / wall,
                note="means over the second half of the run, standard errors from ten blocks; sign convention "
                     "of the reference's output: sigma = -sum 3 Q Q / N per chain, so sigma_xy < 0 under "
                     "positive shear and viscosity = -sigma_xy / rate",
                per_rate=out)


def eta0(horizon=1.0e6, seed=7, n=16384, nk=50):
    """Zero-shear viscosity of the C4 chains at rest by Green-Kubo: eta0 = int_0^inf G(t)
    dt (trapezoid over the multi-tau lags), G(t) the stress autocorrelation of the
    equilibrium run; the standard error is over 16 groups of chains.  Steady shear at a
    rate well below 1/tau_d gives sigma_xy = -rate * eta0 (reference sign convention)."""
    groups = 16
    with eng.Engine(0) as e:
        ens = e.add_ensemble(n, nk, 1.0, 1, seed=seed, sample="gt", delta=1.0, corr_p=16, corr_m=2, corr_levels=18)
        e.init_equilibrium(ens)
        t0 = time.perf_counter()
        jumps, t = 0, 0.0
        while t < horizon:
            t = min(horizon, t + horizon / 10.0)
            jumps += e.advance(ens, t)["events"]
            print("eta0: t = %g tau_K, %.3g jumps, %.1f s" % (t, jumps, time.perf_counter() - t0), file=sys.stderr,
                  flush=True)
        wall = time.perf_counter() - t0
        lag, g = e.correlator(ens)
        _, gg = e.correlator_chains(ens, 0, n, n // groups)
    keep = lag <= horizon / 10.0  # lags with at least ten independent windows behind them
    trap = lambda y: float(np.sum(0.5 * (y[1:] + y[:-1]) * np.diff(lag[keep])))
    per = np.array([trap(x[keep]) for x in gg])
    return dict(workload="eta0: %d chains, Nk=%d, beta=1, CD on, G(t) at rest" % (n, nk), horizon_tau_k=horizon,
                jumps=jumps, seconds=wall, G0=float(g[0]), eta0=trap(g[keep]),
                eta0_se=float(per.std(ddof=1) / np.sqrt(groups)), max_lag=float(lag[keep].max()),
                G_at_max_lag_over_G0=float(g[keep][-1] / g[0]))


def c5(horizon=20.0, window=5.0, seed=1, n_requests=2000, arrival_rate=20.0, graph=0):
    """Serving stream: 2000 requests at 20/s, Nk 16..512, beta {0.5,1,2}, 1024..65536
    chains, half sheared; `horizon` tau_K per request (scaled; see the note)."""
    rep, lat = eng.serve(horizon=horizon, window=window, seed=seed, n_requests=n_requests,
                         arrival_rate=arrival_rate, progress_s=10.0, graph=graph)
    plan = eng.serve_plan(horizon=horizon, seed=seed, n_requests=n_requests, arrival_rate=arrival_rate)
    rep.update(workload="c5: serving stream, %d requests at %g/s" % (n_requests, arrival_rate),
               horizon_tau_k=horizon, window_tau_k=window,
               offered_chains=int(plan["n_chains"].sum()),
               device_busy=rep["advance_s"] / rep["makespan_s"],
               note="scaled horizon: the stated 1e5 tau_K per request is about 1.7e14 jumps in total; "
                    "device_busy = time in KMC rounds / makespan (near 1: the device, not the arrivals, "
                    "sets the pace)")
    return rep


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("which", choices=["c1", "c3", "c4", "c5", "eta0"])
    ap.add_argument("--horizon", type=float, default=None)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--requests", type=int, default=None, help="c5: number of requests")
    ap.add_argument("--window", type=int, default=None, help="c1: samples per staging window")
    ap.add_argument("--graph", type=int, default=None, help="c1, c5: 1 = windows as captured HIP graphs")
    ap.add_argument("--distr", default=None, help="c3: directory for the distr_Z/N/Q.dat tables")
    args = ap.parse_args()
    fn = dict(c1=c1, c3=c3, c4=c4, c5=c5, eta0=eta0)[args.which]
    kw = dict(seed=args.seed)
    if args.horizon is not None:
        kw["horizon"] = args.horizon
    if args.requests is not None and args.which == "c5":
        kw["n_requests"] = args.requests
    if args.window is not None and args.which == "c1":
        kw["window"] = args.window
    if args.graph is not None and args.which in ("c1", "c5"):
        kw["graph"] = args.graph
    if args.distr is not None and args.which == "c3":
        kw["distr"] = args.distr
    print(json.dumps(fn(**kw)), flush=True)


if __name__ == "__main__":
    main()
