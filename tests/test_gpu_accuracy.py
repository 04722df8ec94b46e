"""Error of the device path's G(t) against the FP64 oracle.

The kernels compute rates and strand vectors in FP32 (the FP64 clock and correlator
aside).  The FP64 oracle runs the same chains with the same Philox streams in double
precision; trajectories coincide until a pick lands within FP32 rounding of an event
boundary, so the difference in G(t) measures what FP32 arithmetic does to the result.
Bar: max over lags of |G_fp32 - G_fp64| / G_fp64(0) below 2%.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

CASES = [
    dict(n=1024, nk=25, cd=0, T=200.0),
    dict(n=512, nk=100, cd=1, T=60.0),
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: "nk%d_cd%d" % (c["nk"], c["cd"]))
def test_gt_error_against_fp64_oracle(engine, oracle, case):
    n, nk, cd, T = case["n"], case["nk"], case["cd"], case["T"]
    p, m = 16, 2
    s = oracle.corr_levels(T, 1.0, p, m)
    ens = engine.add_ensemble(n, nk, 1.0, cd, seed=21, sample="gt", delta=1.0, corr_p=p, corr_m=m,
                              corr_levels=s, window=64)
    engine.init_equilibrium(ens)
    engine.advance(ens, T)
    lag, g32 = engine.correlator(ens)
    engine.remove_ensemble(ens)
    P = oracle.Params(1.0, nk, cd)
    O = oracle.Ensemble(P, seed=21, n=n, fp64=True)
    _, samples = O.advance(t_end=T, delta=1.0, k0=0, k_end=int(T), samples=True)
    corr = oracle.Correlator(p, m, s, n)
    for row in samples:
        corr.add_row(row)
    lag64, g64 = corr.result()
    assert np.array_equal(lag, lag64)
    have = np.nonzero(g64 != 0)[0]
    err = float(np.max(np.abs(g32[have] - g64[have])) / g64[0])
    print("G(t) FP32 device vs FP64 oracle, %s: max |dG|/G(0) = %.3g over %d lags" % (
        "nk%d cd%d" % (nk, cd), err, have.size))
    assert err < 0.02


@pytest.mark.parametrize("rate", [1e-3, 1e-2, 1e-1])
def test_shear_stress_error_against_fp64_oracle(engine, oracle, rate):
    """Steady-shear observables of the sweep (shear stress and first normal stress
    difference, time-averaged over the second half of the run): the device's FP32 flow
    path against the FP64 oracle on the same chains and Philox streams.
    Bar: |d sigma_xy| and |d N1| below 2% of the FP64 value's magnitude (+ a floor of
    2% of G(0)-scale stress, <Z>, for rates where the signal is small)."""
    n, nk, T, dlt = 256, 50, 100.0, 1.0
    kappa = [0, rate, 0, 0, 0, 0, 0, 0, 0]
    ens = engine.add_ensemble(n, nk, 1.0, 1, kappa=kappa, seed=31, sample="flow", delta=dlt, window=32)
    engine.init_equilibrium(ens)
    engine.advance(ens, T)
    rows = engine.flow_stress(ens)
    engine.remove_ensemble(ens)
    P = oracle.Params(1.0, nk, 1, kappa=kappa)
    O = oracle.Ensemble(P, seed=31, n=n, fp64=True)
    k_end = int(T / dlt) + 1
    _, s = O.advance(t_end=T, delta=dlt, k0=1, k_end=k_end, samples=True, flow_samples=True)
    ref = s.sum(axis=1)                       # [row][8] sums over chains, FP64
    half = slice(rows.shape[0] // 2, None)
    sxy_g = np.mean(rows[half, 3] / rows[half, 8])
    n1_g = np.mean((rows[half, 0] - rows[half, 1]) / rows[half, 8])
    sxy_o = np.mean(ref[half, 3] / n)
    n1_o = np.mean((ref[half, 0] - ref[half, 1]) / n)
    zbar = np.mean(ref[half, 7] / n)
    e_sxy = abs(sxy_g - sxy_o) / max(abs(sxy_o), zbar)
    e_n1 = abs(n1_g - n1_o) / max(abs(n1_o), zbar)
    print("shear %g: sigma_xy %.5g (FP64 %.5g), N1 %.5g (FP64 %.5g), rel err %.2g / %.2g" % (
        rate, sxy_g, sxy_o, n1_g, n1_o, e_sxy, e_n1))
    assert np.all(rows[:, 8] == n)
    assert e_sxy < 0.02 and e_n1 < 0.02


def _gt_error_report(lag, g32, g64, what):
    """Paired per-chain differences: dG = mean_c (G32_c - G64_c) with its standard error
    (chains are independent).  Over the lags where G64 >= 1% of G64(0): the largest
    |dG|/G64, and whether every |dG| lies within 2% of G64 plus three standard errors."""
    d = g32 - g64
    G = g64.mean(axis=0)
    dG = d.mean(axis=0)
    se = d.std(axis=0, ddof=1) / np.sqrt(d.shape[0])
    sel = np.nonzero(G >= 0.01 * G[0])[0]
    rel = np.abs(dG[sel]) / G[sel]
    ok = np.abs(dG[sel]) <= 0.02 * G[sel] + 3.0 * se[sel]
    resolved = sel[1.96 * se[sel] < 0.02 * G[sel]]
    print("%s: %d lags with G >= 1%% G(0) (up to lag %g); max |dG|/G %.3g at lag %g; max |dG|/G(0) %.3g; "
          "lags resolved at 2%% (95%% CI) %d, max |dG|/G there %.3g" % (
              what, sel.size, lag[sel].max(), rel.max(), lag[sel][np.argmax(rel)],
              np.max(np.abs(dG[sel])) / G[0], resolved.size,
              (np.abs(dG[resolved]) / G[resolved]).max() if resolved.size else float("nan")))
    return ok, sel


@pytest.mark.parametrize("case", [
    dict(name="C1 full horizon", n=1000, nk=25, cd=0, T=1.0e5),
    dict(name="C2 subset", n=512, nk=100, cd=1, T=1.0e4),
], ids=lambda c: c["name"].replace(" ", "_"))
def test_gt_error_at_config_horizon(engine, oracle, case):
    """G(t) of the device (FP32 rates and strand vectors, FP64 clock and correlator)
    against the FP64 oracle on the same chains and Philox streams, at the configuration's
    own horizon: C1 in full (1000 chains, Nk 25, CD off, 1e5 tau_K, 14 levels) and the
    first 512 chains of C2 (Nk 100, CD on) to 1e4 tau_K.  Trajectories part after the
    first pick that lands within FP32 rounding of an event boundary, so the comparison is
    statistical: every lag with G >= 1% of G(0) must agree within 2% of G plus three
    standard errors of the paired per-chain difference."""
    n, nk, cd, T, p, m = case["n"], case["nk"], case["cd"], case["T"], 16, 2
    s = oracle.corr_levels(T, 1.0, p, m)
    ens = engine.add_ensemble(n, nk, 1.0, cd, seed=1, sample="gt", delta=1.0, corr_p=p, corr_m=m, corr_levels=s)
    engine.init_equilibrium(ens)
    engine.advance(ens, T)
    lag, g32 = engine.correlator_chains(ens)
    engine.remove_ensemble(ens)
    lag64, g64, ev, sec = oracle.gt_chains(oracle.Params(1.0, nk, cd), 1, n, T, s=s, fp64=True)
    assert np.array_equal(lag, lag64)
    ok, sel = _gt_error_report(lag, g32, g64, case["name"])
    assert sel.size >= 20
    assert np.all(ok), "lags outside the bar: %s" % lag[sel][~ok]
