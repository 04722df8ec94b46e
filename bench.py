"""Benchmark: KMC jumps/s of the DSM ensemble path on one MI355X, batched vs plain, with
the error of G(t) against the FP64 oracle.

Headline workload (the largest single-GPU configuration): polydisperse melt of 262,144
chains, Nk = round(LogNormal) with number average 200 and Mw/Mn = 1.5 clipped to
[10, 2000], beta = 1, constraint dynamics on with CD partners drawn from the melt's own
length distribution (weight-averaged), G(t) sampled every tau_K into the multi-tau
correlator (16 points per level, block factor 2).  One "step" advances every chain by
--window Kuhn times through the batched persistent form (chains split by Z into six
bins whose kernels run side by side on four streams), plus the correlator update.  All inputs
are synthetic (Philox-keyed equilibrium draw); the state is resident in HBM when the
timed region starts.

Printed: one JSON line.
  value        whole-job jumps/s of the headline workload (wall clock over the timed steps)
  roofline     the KMC pass of one step (six geometry kernels, HIP events from the split
               to the end of the late round): bound "valu" -- the kernels keep each chain
               in registers for the whole window, so HBM sees one read and one write of a
               chain per window and VALU issue is the limit (rocprofv3 counters).
               achieved = VALU wave-instructions per jump (committed rocprofv3 PMC summary
               of the same pass, profiles/) x jumps per step / pass time per step, against
               the issue peak of 256 CUs x 4 SIMDs x 2.4 GHz / 4 cycles per wave64
               instruction.  hbm_byte_model: the algorithmic bytes 20 Z + 32 per jump (a full
               re-read of the chain per jump) over the same time against 8 TB/s;
               hbm_measured: the pass's HBM bytes (PMC, FETCH_SIZE/WRITE_SIZE calibrated on
               a copy) over the same time.
  c2           the monodisperse melt (131,072 chains, Nk 100) as a secondary line.
  plain_vs_batched  the 16-ensemble shear sweep (16 x 16,384 chains, Nk 50): plain = each
               ensemble in its own launch sequence, state in HBM between launches
               (1 jump per chain per launch, the reference's structure, and 16 per launch);
               batched = all ensembles in one persistent pass per window.
  cpu_baseline the CPU oracle (same algorithm, same Philox streams, G(t) sampling) timed on
               this host on a bounded chain subset of the headline workload: all threads
               (value), one thread, and the FP64 oracle.
  gt_err_vs_fp64  1000 chains, Nk 25, CD off, to 1e5 tau_K (the small configuration in full):
               device G(t) against the FP64 oracle, paired per-chain differences with
               standard errors, over the lags where G >= 1% of G(0).
  stress_err_vs_fp64  shear from equilibrium of the sweep's chains (512 chains, Nk 50, CD on,
               500 tau_K, rates 1e-3, 1e-2, 1e-1): means of sigma_xy and N1 over the second
               half (strain 0.25-0.5, 2.5-5, 25-50: start-up, not yet steady at the lower
               rates), device against the FP64 oracle on the same chains and Philox streams.
  dist_err_vs_fp64  the Z / N / Q distributions of the first 4096 chains of the headline
               ensemble (own melt partners) after 100 tau_K: device against the FP64
               oracle on the same chains and Philox streams (largest pmf / CDF
               differences), and each against the equilibrium laws in sampling-error units.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK_HBM_GBS = 8000.0
# VALU issue peak of MI355X: 256 CUs x 4 SIMDs, one wave64 instruction per 4 cycles at 2.4 GHz
N_SIMD, CLOCK_GHZ = 1024, 2.4
PEAK_VALU_G = N_SIMD * CLOCK_GHZ / 4.0  # G wave-instructions per second
C3 = dict(n_chains=262144, beta=1.0, cd=1, corr_p=16, corr_m=2, delta=1.0)
C2 = dict(n_chains=131072, nk=100, beta=1.0, cd=1, corr_p=16, corr_m=2, delta=1.0)
C4 = dict(n_ens=16, n_chains=16384, nk=50, beta=1.0, cd=1)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--window", type=int, default=16, help="Kuhn times (samples) per step")
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--only", choices=["c3", "c2"], default=None, help="run one workload (profiling)")
    ap.add_argument("--record", default=None, help="write the jumps of every advance call to this JSON file")
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline and the FP64 G(t) error")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=8.0, help="target duration of the all-thread CPU run")
    return ap.parse_args()


def counters(name):
    """Per-jump figures of a KMC pass from a committed rocprofv3 PMC summary (profiles/)."""
    path = os.path.join(ROOT, "profiles", name)
    try:
        return json.load(open(path))
    except Exception:
        return {}


def timed_steps(eng, ids, args, record):
    """Warm-up and timed steps of --window tau_K over the given ensembles (one batched pass
    per step); returns totals of the timed steps."""
    t = 0.0
    for _ in range(args.warmup):
        t += args.window
        st = eng.advance_many(ids, t)
        record.append(st["events"])
    ev = zs = 0
    kmc_ms = first_ms = 0.0
    t0 = time.perf_counter()
    for _ in range(args.steps):
        t += args.window
        st = eng.advance_many(ids, t)
        record.append(st["events"])
        ev += st["events"]
        zs += st["z_sum"]
        kmc_ms += st["kmc_ms"]
        first_ms += st["first_kmc_ms"]
    wall = time.perf_counter() - t0
    return dict(jumps=ev, z_sum=zs, wall_s=wall, kmc_ms=kmc_ms, first_ms=first_ms, t_end=t)


def roofline(tot, steps, cnt, pass_ms):
    """VALU roofline of the pass (bound per the counters), with both HBM views."""
    jumps_step = tot["jumps"] / steps
    bytes_model = (20.0 * tot["z_sum"] + 32.0 * tot["jumps"]) / steps
    valu_pj = cnt.get("valu_per_jump")
    hbm_pj = cnt.get("hbm_bytes_per_jump")
    achieved = valu_pj * jumps_step / (pass_ms * 1e-3) / 1e9 if valu_pj else None
    model_gbs = bytes_model / (pass_ms * 1e-3) / 1e9
    meas = hbm_pj * jumps_step if hbm_pj else None
    return {"bound": "valu", "achieved": achieved, "peak": PEAK_VALU_G, "unit": "G VALU wave-instructions/s",
            "frac": achieved / PEAK_VALU_G if achieved else None, "traffic": meas,
            "valu_per_jump": valu_pj, "waves_per_simd": cnt.get("waves_per_simd"),
            "pass_ms": pass_ms, "jumps_per_pass": jumps_step,
            "counters": cnt.get("source"),
            "hbm_byte_model": {"bytes_per_pass": bytes_model, "achieved_GBs": model_gbs, "peak_GBs": PEAK_HBM_GBS,
                               "frac": model_gbs / PEAK_HBM_GBS,
                               "note": "algorithmic 20 Z + 32 bytes per jump (a full re-read of the chain per jump)"},
            "hbm_measured": {"bytes_per_pass": meas,
                             "achieved_GBs": meas / (pass_ms * 1e-3) / 1e9 if meas else None,
                             "frac": meas / (pass_ms * 1e-3) / 1e9 / PEAK_HBM_GBS if meas else None}}


def run_c3(g, args, record):
    from gpu_dsm_amd.workloads import c3_lengths
    nk = c3_lengths(C3["n_chains"], seed=args.seed)
    T_total = float(args.window * (args.warmup + args.steps + 1))
    levels = max(2, int(math.ceil(math.log(T_total / C3["corr_p"]) / math.log(C3["corr_m"]))) + 1)
    with g.Engine(0) as eng:
        ens = eng.add_ensemble(C3["n_chains"], int(nk.max()), C3["beta"], C3["cd"], seed=args.seed, sample="gt",
                               delta=C3["delta"], corr_p=C3["corr_p"], corr_m=C3["corr_m"], corr_levels=levels,
                               window=args.window, nk_per_chain=nk, partner="melt")
        eng.init_equilibrium(ens)
        tot = timed_steps(eng, [ens], args, record)
        lag, gt = eng.correlator(ens)
        z = eng.check(ens, t_min=tot["t_end"])
    tot["G0"] = float(gt[0])
    tot["violations"] = int(sum(z.values()))
    tot["nk"] = nk
    return tot


def run_c2(g, args, record):
    T_total = float(args.window * (args.warmup + args.steps + 1))
    levels = max(2, int(math.ceil(math.log(T_total / C2["corr_p"]) / math.log(C2["corr_m"]))) + 1)
    with g.Engine(0) as eng:
        ens = eng.add_ensemble(C2["n_chains"], C2["nk"], C2["beta"], C2["cd"], seed=args.seed, sample="gt",
                               delta=C2["delta"], corr_p=C2["corr_p"], corr_m=C2["corr_m"], corr_levels=levels,
                               window=args.window)
        eng.init_equilibrium(ens)
        tot = timed_steps(eng, [ens], args, record)
        lag, gt = eng.correlator(ens)
    tot["G0"] = float(gt[0])
    return tot


def run_plain_vs_batched(g, args):
    """C4 shear sweep: the same 16 ensembles and horizon in three forms, fresh engines."""
    rates = [10.0 ** (-5.0 + 4.0 * q / (C4["n_ens"] - 1)) for q in range(C4["n_ens"])]
    T = 8.0
    out = {}
    for form, jumps in (("batched", 0), ("plain_16", 16), ("plain_1", 1)):
        with g.Engine(0) as eng:
            if jumps:
                eng.set_option("plain_jumps", jumps)
            ids = []
            for q, gd in enumerate(rates):
                ens = eng.add_ensemble(C4["n_chains"], C4["nk"], C4["beta"], C4["cd"],
                                       kappa=[0, gd, 0, 0, 0, 0, 0, 0, 0], seed=args.seed, salt=q, sample="flow",
                                       delta=1.0, window=16)
                eng.init_equilibrium(ens)
                ids.append(ens)
            ev = 0
            t0 = time.perf_counter()
            if form == "batched":
                ev = eng.advance_many(ids, T)["events"]
            else:
                for ens in ids:  # one ensemble per launch sequence
                    ev += eng.advance(ens, T, mode="plain")["events"]
            wall = time.perf_counter() - t0
        out[form] = {"jumps": ev, "seconds": wall, "jumps_per_s": ev / wall}
    b = out["batched"]["jumps_per_s"]
    return {"workload": "16 ensembles x 16384 chains, Nk=50, beta=1, CD on, shear 1e-5..1e-1, %g tau_K" % T,
            "batched": out["batched"], "plain_16_jumps_per_launch": out["plain_16"],
            "plain_1_jump_per_launch": out["plain_1"],
            "batched_over_plain_16": b / out["plain_16"]["jumps_per_s"],
            "batched_over_plain_1": b / out["plain_1"]["jumps_per_s"]}


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except Exception:
        pass
    return None


def cpu_threads():
    try:
        n = len(os.sched_getaffinity(0))
    except Exception:
        n = os.cpu_count() or 1
    return max(1, min(n, 16))  # a one-GPU box's CPU share is 16 threads


def cpu_baseline(args, nk_full):
    """The oracle on the first chains of the headline ensemble (same lengths, same melt
    partner tables, same Philox streams), G(t) sampled every tau_K: all threads, one
    thread, and FP64."""
    from oracle import oracle as O
    nmax = int(nk_full.max())
    P = O.Params(C3["beta"], nmax, C3["cd"], nmax=nmax, partner="melt", nk_chains=nk_full)
    threads = cpu_threads()

    def run(n, T, th, fp64):
        lags, G, ev, sec = O.gt_chains(P, args.seed, n, T, s=max(1, O.corr_levels(T, 1.0, 16, 2)), threads=th,
                                       fp64=fp64, nk_per_chain=nk_full[:n])
        return ev, sec

    ev, sec = run(256, 8.0, threads, False)  # calibration
    rate_chain_tau = ev / (256 * 8.0)
    per_s = ev / max(sec, 1e-9)
    n = 4096
    T = float(max(8.0, round(args.cpu_seconds * per_s / (n * rate_chain_tau))))
    ev, sec = run(n, T, threads, False)
    all_rate = ev / sec
    n1 = 256
    T1 = float(max(4.0, round(0.5 * args.cpu_seconds * (per_s / threads) / (n1 * rate_chain_tau))))
    ev1, sec1 = run(n1, T1, 1, False)
    ev64, sec64 = run(n1 * threads, T1, threads, True)
    return {"value": all_rate, "unit": "jumps/s", "cores": threads, "kind": "port",
            "sample": "first %d chains of the headline ensemble (same lengths, melt partner tables and Philox "
                      "streams) from equilibrium to t=%g tau_K with G(t) sampling, mirror-FP32 oracle on %d host "
                      "threads: %d jumps in %.1f s" % (n, T, threads, ev, sec),
            "cpu_model": cpu_model(), "host_threads_available": os.cpu_count(),
            "single_thread": {"value": ev1 / sec1, "sample": "%d chains to t=%g tau_K: %d jumps in %.1f s"
                                                             % (n1, T1, ev1, sec1)},
            "fp64_oracle": {"value": ev64 / sec64, "cores": threads,
                            "sample": "%d chains to t=%g tau_K: %d jumps in %.1f s" % (n1 * threads, T1, ev64,
                                                                                      sec64)}}


def gt_error(g, args):
    """The small configuration in full: 1000 chains, Nk 25, CD off, 1e5 tau_K, p 16, m 2."""
    from oracle import oracle as O
    n, nk, T, p, m = 1000, 25, 1.0e5, 16, 2
    s = O.corr_levels(T, 1.0, p, m)
    with g.Engine(0) as eng:
        ens = eng.add_ensemble(n, nk, 1.0, 0, seed=args.seed, sample="gt", delta=1.0, corr_p=p, corr_m=m,
                               corr_levels=s)
        eng.init_equilibrium(ens)
        t0 = time.perf_counter()
        st = eng.advance(ens, T)
        gpu_s = time.perf_counter() - t0
        lag, g32 = eng.correlator_chains(ens)
    lag64, g64, ev64, sec64 = O.gt_chains(O.Params(1.0, nk, 0), args.seed, n, T, s=s, fp64=True,
                                          threads=cpu_threads())
    d = g32 - g64
    G = g64.mean(axis=0)
    dG = d.mean(axis=0)
    se = d.std(axis=0, ddof=1) / math.sqrt(n)
    sel = np.nonzero(G >= 0.01 * G[0])[0]
    rel = np.abs(dG[sel]) / G[sel]
    within = np.abs(dG[sel]) <= 0.02 * G[sel] + 3.0 * se[sel]
    resolved = sel[1.96 * se[sel] < 0.02 * G[sel]]
    k = int(np.argmax(rel))
    return {"workload": "1000 chains, Nk=25, beta=1, CD off, G(t) to 1e5 tau_K (p 16, m 2, %d levels)" % s,
            "lags_compared": int(sel.size), "max_lag": float(lag[sel].max()),
            "max_abs_err_over_G0": float(np.max(np.abs(dG[sel])) / G[0]),
            "max_rel_err": float(rel[k]), "at_lag": float(lag[sel][k]),
            "stderr_at_that_lag_rel": float(se[sel][k] / G[sel][k]),
            "lags_within_2pct_plus_3se": int(within.sum()),
            "lags_resolved_at_2pct_95ci": int(resolved.size),
            "max_rel_err_resolved": float((np.abs(dG[resolved]) / G[resolved]).max()) if resolved.size else None,
            "gpu_jumps": int(st["events"]), "gpu_seconds": gpu_s, "fp64_oracle_jumps": ev64,
            "fp64_oracle_seconds": sec64}


def stress_error(g, args, n=512, nk=50, T=500.0):
    """Shear stress of the same chains from equilibrium: device (FP32 flow path) vs FP64
    oracle, second-half means."""
    from oracle import oracle as O
    out = {"workload": "%d chains, Nk=%d, beta=1, CD on, shear, %g tau_K, second-half means" % (n, nk, T),
           "rates": []}
    for rate in (1e-3, 1e-2, 1e-1):
        kappa = [0, rate, 0, 0, 0, 0, 0, 0, 0]
        with g.Engine(0) as eng:
            ens = eng.add_ensemble(n, nk, 1.0, 1, kappa=kappa, seed=args.seed + 30, sample="flow", delta=1.0,
                                   window=64)
            eng.init_equilibrium(ens)
            eng.advance(ens, T)
            rows = eng.flow_stress(ens)
        E = O.Ensemble(O.Params(1.0, nk, 1, kappa=kappa), seed=args.seed + 30, n=n, fp64=True)
        _, smp = E.advance(t_end=T, delta=1.0, k0=1, k_end=int(T) + 1, samples=True, flow_samples=True,
                           threads=cpu_threads())
        ref = smp.sum(axis=1)
        half = slice(rows.shape[0] // 2, None)
        sxy, sxy64 = float(np.mean(rows[half, 3] / rows[half, 8])), float(np.mean(ref[half, 3] / n))
        n1, n164 = (float(np.mean((rows[half, 0] - rows[half, 1]) / rows[half, 8])),
                    float(np.mean((ref[half, 0] - ref[half, 1]) / n)))
        out["rates"].append({"rate": rate, "sigma_xy": sxy, "sigma_xy_fp64": sxy64,
                             "rel_err_sigma_xy": abs(sxy - sxy64) / abs(sxy64), "N1": n1, "N1_fp64": n164,
                             "rel_err_N1": abs(n1 - n164) / abs(n164)})
    return out


def dist_error(g, args, nk_full, n=4096, T=100.0):
    """Z / N / Q distributions of the same chains after T tau_K: device vs FP64 oracle,
    and each against the equilibrium laws (gpu_dsm_amd.distributions)."""
    from oracle import oracle as O
    from gpu_dsm_amd import distributions as D
    nk = np.ascontiguousarray(nk_full[:n], np.int32)
    nmax = int(nk.max())
    with g.Engine(0) as eng:
        ens = eng.add_ensemble(n, nmax, C3["beta"], C3["cd"], seed=args.seed, sample="none", nk_per_chain=nk,
                               partner="melt")
        eng.init_equilibrium(ens)
        eng.advance(ens, T)
        dev = D.fetch_zn_q(eng, ens)
    P = O.Params(C3["beta"], nmax, C3["cd"], nmax=nmax, partner="melt", nk_chains=nk)
    E = O.Ensemble(P, seed=args.seed, n=n, fp64=True, nk_per_chain=nk)
    t0 = time.perf_counter()
    ev = E.advance(t_end=T, threads=cpu_threads())[0]
    sec = time.perf_counter() - t0
    st = E.state()
    ref = (st["Z"], st["N"], st["Q"])
    out = {"workload": "first %d chains of the headline lengths (own melt partners), beta=1, CD on, %g tau_K"
                       % (n, T)}
    out.update(D.two_sample(dev, ref))
    out["device_vs_equilibrium"] = D.summarize(*dev, nk, C3["beta"])
    out["fp64_oracle_vs_equilibrium"] = D.summarize(*ref, nk, C3["beta"])
    out["fp64_oracle_jumps"] = int(ev)
    out["fp64_oracle_seconds"] = sec
    return out


def main():
    args = parse()
    if args.gpus != 1:
        print("bench.py: the KMC path runs on exactly one GPU; --gpus %d is not supported (across GPUs "
              "it would shard by chain ranges, one engine per GPU, reducing only stress and correlator sums)"
              % args.gpus, file=sys.stderr)
        sys.exit(2)
    import gpu_dsm_amd as g

    record = []
    if args.only == "c2":
        c2 = run_c2(g, args, record)
        if args.record:
            json.dump({"workload": "c2", "advance_jumps": record}, open(args.record, "w"))
        copy_ms = g.probe_copy(1 << 30, 5)  # the byte counters' reference in the same profiled run
        print(json.dumps({"c2_jumps_per_s": c2["jumps"] / c2["wall_s"], "jumps": c2["jumps"],
                          "row_kernel_ms": c2["first_ms"] / args.steps, "copy_ms": copy_ms}))
        return
    c3 = run_c3(g, args, record)
    if args.record:
        json.dump({"workload": "c3", "advance_jumps": record}, open(args.record, "w"))
    if args.only == "c3":
        copy_ms = g.probe_copy(1 << 30, 5)  # the byte counters' reference in the same profiled run
        print(json.dumps({"c3_jumps_per_s": c3["jumps"] / c3["wall_s"], "jumps": c3["jumps"],
                          "pass_ms": c3["kmc_ms"] / args.steps, "copy_ms": copy_ms}))
        return
    value = c3["jumps"] / c3["wall_s"]
    pass_ms = c3["kmc_ms"] / args.steps
    rl = roofline(c3, args.steps, counters("kmc_pass_c3_counters.json"), pass_ms)
    c2 = run_c2(g, args, [])
    rl2 = roofline(c2, args.steps, counters("kmc_row_form_c2_counters.json"), c2["first_ms"] / args.steps)
    pvb = None if args.no_plain else run_plain_vs_batched(g, args)
    copy_bytes = 1 << 30
    copy_ms = g.probe_copy(copy_bytes, 20)
    cpu = gte = dse = sse = None
    if not args.no_cpu:
        try:
            cpu = cpu_baseline(args, c3["nk"])
        except Exception as exc:  # the baseline is reported, never required
            cpu = {"value": None, "unit": "jumps/s", "cores": 0, "kind": "port", "sample": "failed: %s" % exc}
        try:
            gte = gt_error(g, args)
        except Exception as exc:
            gte = {"failed": str(exc)}
        try:
            dse = dist_error(g, args, c3["nk"])
        except Exception as exc:
            dse = {"failed": str(exc)}
        try:
            sse = stress_error(g, args)
        except Exception as exc:
            sse = {"failed": str(exc)}
    out = {
        "metric": "KMC jumps/sec & G(t) err vs FP64 oracle, 1 MI355X, batched vs plain",
        "value": value,
        "unit": "jumps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * c3["wall_s"] / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (Philox-keyed DSM equilibrium draw)",
        "config": {"workload": "polydisperse melt, 262144 chains, Nk log-normal (Mn 200, PDI 1.5) clipped to "
                               "[10, 2000], beta=1, CD on with melt partners, G(t) multi-tau p=16 m=2, "
                               "%d tau_K per step" % args.window,
                   "form": "batched persistent: chains split by Z into six bins (row form 4 chains per wave "
                           "Z<64, two-group row form 4 per wave Z<128, two-group pair form 2 per wave Z<240, "
                           "wave form 1 per wave Z<480/960/2048) run side by side on four streams, then a late round for "
                           "chains that outgrew theirs",
                   "n_chains": C3["n_chains"], "nk_max": int(c3["nk"].max()), "nk_mean": float(c3["nk"].mean()),
                   "window_tau_k": args.window},
        "roofline": rl,
        "cpu_baseline": cpu,
        "gt_err_vs_fp64": gte,
        "dist_err_vs_fp64": dse,
        "stress_err_vs_fp64": sse,
        "jumps": c3["jumps"],
        "mean_Z_per_jump": c3["z_sum"] / max(c3["jumps"], 1),
        "kmc_device_ms": c3["kmc_ms"],
        "G0": c3["G0"],
        "invariant_violations": c3["violations"],
        "c2": {"workload": "monodisperse melt, 131072 chains, Nk=100, beta=1, CD on, G(t) p=16 m=2, %d tau_K per "
                           "step" % args.window,
               "jumps_per_s": c2["jumps"] / c2["wall_s"], "ms_per_step": 1e3 * c2["wall_s"] / args.steps,
               "row_kernel_avg_ms": c2["first_ms"] / args.steps, "roofline_row_kernel": rl2, "G0": c2["G0"]},
        "plain_vs_batched": pvb,
        "copy_probe": {"bytes": copy_bytes, "ms": copy_ms, "GBs": 2.0 * copy_bytes / (copy_ms * 1e-3) / 1e9,
                       "note": "plain 16-byte device copy, read+write bytes over time: achievable HBM rate"},
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
