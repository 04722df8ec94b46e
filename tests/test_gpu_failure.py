"""Failure paths of the device path, forced from the host so every kernel exits normally.

  * Stalls (common/ensemble_kernel.cu:264-270, common/eq_ensemble_kernel.cu:396-401): a
    chain whose waiting time is 0 (an overflowing rate), NaN or inf stops with flag 1, 2
    or 3.  A strong extension drives strand vectors until their sliding rates overflow;
    the flow rows then count each chain up to its stall and not after
    (common/ensemble_call_block.cu:471-491, per output step), exactly as the oracle does.
    Chains given a NaN strand vector stall at their first jump and are left out of G(t).
  * run_flag cleared (common/ensemble_call_block.cu:214,353): the call returns -1 after
    the window in progress, with a consistent state that a later call continues
    bit-exactly.
  * The plain form's dsm_advance reports the jumps it made.
"""
import numpy as np
import pytest

from test_gpu_parity import assert_same_state

pytestmark = pytest.mark.gpu


def test_stall_under_strong_extension(engine, oracle):
    n, nk, T, dlt = 256, 20, 3.0, 0.1
    kappa = [4.0, 0, 0, 0, -2.0, 0, 0, 0, -2.0]
    P = oracle.Params(1.0, nk, 1, kappa=kappa)
    ens = engine.add_ensemble(n, nk, 1.0, 1, kappa=kappa, seed=77, sample="flow", delta=dlt, window=8)
    engine.init_equilibrium(ens)
    O = oracle.Ensemble(P, seed=77, n=n)
    k_end = int(round(T / dlt)) + 1
    # the device repeats the oracle's jumps bit for bit: it runs only once the oracle has
    # shown that every chain stalls or reaches T within a bounded number of jumps
    _, s = O.advance(t_end=T, delta=dlt, k0=1, k_end=k_end, samples=True, flow_samples=True, max_events=2000)
    o = O.state()
    assert np.all((o["stall"] != 0) | (o["t"] >= T)) and o["ev"].max() < 2000
    engine.advance(ens, T)
    g = engine.state(ens)
    assert_same_state(g, o, "strong extension")
    assert np.array_equal(g["stall"], o["stall"])
    stalled = np.nonzero(o["stall"])[0]
    assert stalled.size >= 5, "the extension should stall some chains (%d)" % stalled.size
    assert set(np.unique(o["stall"][stalled])) <= {1, 2, 3}
    rows = engine.flow_stress(ens)
    ok = (s[:, :, 7] > 0) & np.all(np.isfinite(s), axis=2)
    want = np.where(ok[..., None], s.astype(np.float64), 0.0).sum(axis=1)
    np.testing.assert_allclose(rows[:, :8], want, rtol=1e-12, atol=1e-9)
    np.testing.assert_array_equal(rows[:, 8], ok.sum(axis=1))
    # the healthy count falls over time, and stalled chains count before their stall
    assert rows[0, 8] > rows[-1, 8]
    assert np.any(ok[:, stalled].any(axis=0))
    engine.remove_ensemble(ens)


def test_nan_chains_stall_and_leave_gt(engine, oracle):
    n, nk, T, p, m = 200, 25, 50.0, 16, 2
    s = oracle.corr_levels(T, 1.0, p, m)
    P = oracle.Params(1.0, nk, 0)
    ens = engine.add_ensemble(n, nk, 1.0, 0, seed=12, sample="gt", delta=1.0, corr_p=p, corr_m=m, corr_levels=s,
                              window=32)
    engine.init_equilibrium(ens)
    st = engine.state(ens)
    bad = np.array([3, 50, 51, 199])
    bad = bad[st["Z"][bad] >= 3]
    assert bad.size >= 2
    st["Q"][bad, 1, 0] = np.nan
    engine.set_state(ens, st["Z"], st["N"], st["Q"], st["w"], t=st["t"], ev=st["ev"], stall=st["stall"],
                     dt_prev=st["dt_prev"], next_sample=st["next_sample"])
    engine.advance(ens, T)
    g = engine.state(ens)
    assert np.all(g["stall"][bad] == 2) and np.all(g["ev"][bad] == 0)
    good = np.setdiff1d(np.arange(n), bad)
    assert np.all(g["stall"][good] == 0)
    O = oracle.Ensemble(P, seed=12, n=n)
    _, samples = O.advance(t_end=T, delta=1.0, k0=0, k_end=int(T), samples=True)
    o = O.state()
    assert np.array_equal(g["Z"][good], o["Z"][good]) and np.array_equal(g["t"][good], o["t"][good])
    corr = oracle.Correlator(p, m, s, good.size)
    for row in samples:
        corr.add_row(row[good])
    lag_o, g_o = corr.result()
    lag, gt = engine.correlator(ens)
    assert np.array_equal(lag, lag_o)
    np.testing.assert_allclose(gt, g_o, rtol=1e-10, atol=1e-12)
    engine.remove_ensemble(ens)


@pytest.mark.parametrize("sample", ["gt", "flow"])
def test_run_flag_abort_then_continue(engine, oracle, sample):
    """A call with the flag cleared stops after its first window (8 samples to t = 8):
    every chain stands at its first jump past t = 8, as in an uninterrupted run, and a
    later call continues to T bit-exactly."""
    n, nk, T = 300, 40, 30.0
    kw = dict(sample=sample, delta=1.0, window=8)
    if sample == "gt":
        kw.update(corr_p=16, corr_m=2, corr_levels=3)
    ens = engine.add_ensemble(n, nk, 1.0, 1, seed=19, **kw)
    engine.init_equilibrium(ens)
    st = engine.advance(ens, T, run_flag=0)
    assert st.get("aborted"), "a cleared run flag must abort the call"
    assert 0 < st["events"]
    P = oracle.Params(1.0, nk, 1)
    O = oracle.Ensemble(P, seed=19, n=n)
    ev_mid, _ = O.advance(t_end=8.0)
    assert st["events"] == ev_mid
    assert_same_state(engine.state(ens), O.state(), "after the abort")
    engine.advance(ens, T)
    O.advance(t_end=T)
    assert_same_state(engine.state(ens), O.state(), "continued after abort")
    if sample == "flow":
        assert engine.flow_stress(ens).shape[0] == int(T)
    engine.remove_ensemble(ens)


def test_plain_advance_reports_its_jumps(engine, oracle):
    n, nk, T = 200, 30, 6.0
    ens = engine.add_ensemble(n, nk, 1.0, 1, seed=23)
    engine.init_equilibrium(ens)
    st = engine.advance(ens, T, mode="plain")
    O = oracle.Ensemble(oracle.Params(1.0, nk, 1), seed=23, n=n)
    ev, _ = O.advance(t_end=T)
    assert st["events"] == ev > 0
    assert_same_state(engine.state(ens), O.state(), "plain advance")
    engine.remove_ensemble(ens)
