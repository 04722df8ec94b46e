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


MIX = [  # (chains, Nk, beta, shear rate, admitted before round)
    (300, 40, 1.0, 0.0, 0),
    (200, 64, 0.5, 1e-2, 0),
    (100, 200, 2.0, 0.0, 1),     # Z crosses 64: row form hands chains to the wave form
    (150, 500, 0.5, 3e-3, 1),    # Z ~ 330: wave form, two 256-position groups
    (250, 30, 1.0, 0.0, 3),      # admitted into the ranges freed by the first two
]


def _add(e, q, n, nk, beta, rate, horizon):
    if rate > 0.0:
        return e.add_ensemble(n, nk, beta, 1, kappa=[0, rate, 0, 0, 0, 0, 0, 0, 0], seed=77, salt=q, sample="flow",
                              delta=1.0, window=16)
    return e.add_ensemble(n, nk, beta, 1, seed=77, salt=q, sample="gt", delta=1.0, corr_p=16, corr_m=2,
                          corr_levels=2, window=16)


def _result(e, ens, rate):
    st = e.state(ens)
    out = dict(Z=st["Z"], N=st["N"], Q=st["Q"], w=st["w"], t=st["t"], ev=st["ev"])
    if rate > 0.0:
        out["rows"] = e.flow_stress(ens)
    else:
        out["lag"], out["G"] = e.correlator(ens)
    return out


def test_results_same_alone_and_in_the_mix():
    """Continuous batching changes no result: requests admitted at different rounds,
    advanced together (one batched pass per 5 tau_K window over every active request)
    and removed when done give the same chain state, G(t) and flow stress rows as each
    request run alone in a fresh engine.  Tolerance: none (bit-identical), since every
    chain's random stream is keyed by (seed, request, chain) and the reductions sum
    in a fixed order relative to the request's first chain."""
    import gpu_dsm_amd as g
    horizon, window = 15.0, 5.0
    mixed, active, t = {}, {}, {}
    with g.Engine(0) as e:
        rnd = 0
        while len(mixed) < len(MIX):
            for q, (n, nk, beta, rate, when) in enumerate(MIX):
                if when == rnd:
                    active[q] = _add(e, q, n, nk, beta, rate, horizon)
                    e.init_equilibrium(active[q])
                    t[q] = 0.0
            qs = sorted(active)
            e.advance_many([active[q] for q in qs], [min(horizon, t[q] + window) for q in qs])
            for q in qs:
                t[q] = min(horizon, t[q] + window)
                if t[q] >= horizon:
                    mixed[q] = _result(e, active[q], MIX[q][3])
                    e.remove_ensemble(active.pop(q))
            rnd += 1
            assert rnd < 20
    for q, (n, nk, beta, rate, _) in enumerate(MIX):
        with g.Engine(0) as e:
            ens = _add(e, q, n, nk, beta, rate, horizon)
            e.init_equilibrium(ens)
            for k in range(1, 4):
                e.advance(ens, k * window)
            alone = _result(e, ens, rate)
        for key, v in alone.items():
            assert np.array_equal(v, mixed[q][key]), "request %d: %s differs alone vs in the mix" % (q, key)
        assert int(alone["Z"].max()) > 0
    assert int(mixed[2]["Z"].max()) > 64 or int(mixed[3]["Z"].max()) > 256
