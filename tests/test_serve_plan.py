"""The serving request stream (host only): determinism, ranges and the memory estimate."""
import numpy as np
import pytest


def test_stream_is_deterministic_and_in_range():
    import gpu_dsm_amd as g
    a = g.serve_plan(seed=5)
    b = g.serve_plan(seed=5)
    c = g.serve_plan(seed=6)
    for k in a:
        assert np.array_equal(a[k], b[k])
    assert not np.array_equal(a["arrival_s"], c["arrival_s"])
    n = len(a["arrival_s"])
    assert n == 2000
    assert np.all(np.diff(a["arrival_s"]) > 0)
    # Poisson at 20/s: mean gap 0.05 s within 4 standard errors
    gaps = np.diff(np.concatenate([[0.0], a["arrival_s"]]))
    assert abs(gaps.mean() - 0.05) < 4 * 0.05 / np.sqrt(n)
    assert set(np.unique(a["nk"])) <= set(range(16, 513, 16))
    assert set(np.unique(a["beta"]).tolist()) <= {0.5, 1.0, 2.0}
    assert a["n_chains"].min() >= 1024 and a["n_chains"].max() <= 65536
    # log-uniform chain counts: median near sqrt(1024 * 65536) = 8192
    assert 6000 < np.median(a["n_chains"]) < 11000
    sheared = a["shear_rate"] > 0
    assert abs(sheared.mean() - 0.5) < 4 * 0.5 / np.sqrt(n)
    r = a["shear_rate"][sheared]
    assert r.min() >= 1e-5 and r.max() <= 1e-2


def test_memory_estimate_grows_with_the_request():
    import gpu_dsm_amd as g
    a = g.serve_plan(seed=2, n_requests=400, flow_fraction=1.0)
    per_strand = a["bytes"] / (a["n_chains"] * ((a["nk"] + 3) // 4 * 4))
    assert np.all(per_strand >= 20)          # at least the 20 B strand records
    eq = g.serve_plan(seed=2, n_requests=400, flow_fraction=0.0)
    assert np.all(eq["bytes"] > 0)


@pytest.mark.parametrize("bad", [dict(n_requests=0), dict(nk_max=2048), dict(nk_step=0), dict(arrival_rate=0.0),
                                 dict(beta=()), dict(hbm_fraction=1.5), dict(horizon=0.0)])
def test_invalid_streams_are_rejected(bad):
    import gpu_dsm_amd as g
    with pytest.raises(g.DsmError):
        g.serve_plan(**bad)
