"""A short tuning run on the device: every candidate runs, the result is cached under the
workload's signature, and a second call returns the cached choice."""
import pytest

pytestmark = pytest.mark.gpu


def test_tune_small_workload(tmp_path):
    import gpu_dsm_amd as g
    from gpu_dsm_amd import autotune as A
    descs = [dict(n_chains=2048, nk=40, beta=1.0, cd=1, seed=3, sample="gt", delta=1.0, corr_p=16, corr_m=2,
                  corr_levels=3)]
    cache = str(tmp_path / "tune.json")
    res = A.tune(descs, step_tau=4.0, reps=1, cache=cache)
    assert not res["cached"] and res["default_rate"] > 0 and res["best_rate"] >= res["default_rate"] * 0.999
    assert set(res["best"]) == set(A.DEFAULTS)
    n_trials = 1 + sum(len(v) - 1 for v in A.CANDIDATES.values()) + 1
    assert len(res["trials"]) == n_trials
    again = A.tune(descs, cache=cache)
    assert again["cached"] and again["best"] == res["best"]
    with g.Engine(0) as e:
        A.apply(e, res["best"])
        opts = e.options()
        assert all(opts[k] == v for k, v in res["best"].items())
