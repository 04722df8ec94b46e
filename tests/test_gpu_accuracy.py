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
