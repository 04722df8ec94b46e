"""A short serving stream on the device: every request completes, latencies are sane."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_short_stream_completes():
    import gpu_dsm_amd as g
    kw = dict(n_requests=24, arrival_rate=200.0, seed=3, nk_min=16, nk_max=64, nk_step=16, nch_min=256,
              nch_max=2048, horizon=20.0, window=5.0, hbm_fraction=0.5)
    rep, lat = g.serve(**kw)
    plan = g.serve_plan(**kw)
    assert rep["completed"] == 24
    assert np.all(lat > 0)
    assert rep["p50_s"] <= rep["p99_s"] <= rep["max_s"] + 1e-12
    assert rep["makespan_s"] >= plan["arrival_s"][-1]
    assert rep["jumps"] > 0 and rep["rounds"] >= 4
    assert 1 <= rep["max_active"] <= 24
    print("serving: p50 %.4f s  p99 %.4f s  %.3g jumps/s" % (rep["p50_s"], rep["p99_s"], rep["jumps_per_s"]))


def test_removed_ensembles_are_reused(engine, oracle):
    """Chains, strand slots and staging rows of removed ensembles are handed out again:
    repeated add/run/remove cycles keep the device footprint flat and results exact."""
    import gpu_dsm_amd as g
    from test_gpu_parity import assert_same_state
    free = []
    for k in range(5):
        n, nk = 600 + 37 * k, 40
        ens = engine.add_ensemble(n, nk, 1.0, 1, seed=90 + k, sample="gt", corr_p=16, corr_m=2, window=8)
        engine.init_equilibrium(ens)
        engine.advance_events(ens, 30)
        O = oracle.Ensemble(oracle.Params(1.0, nk, 1), seed=90 + k, n=n)
        O.advance(max_events=30)
        assert_same_state(engine.state(ens), O.state(), "cycle %d" % k)
        engine.remove_ensemble(ens)
        free.append(g.mem_info(0)[0])
    assert free[-1] >= free[1] - (64 << 20)
