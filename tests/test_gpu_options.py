"""Engine options (dsm_set_option) change how passes are laid out, never a result.

  * The plain form with several jumps per launch (state in HBM between launches) stays
    bit-exact against the oracle.
  * A polydisperse ensemble with melt partners (every geometry, concurrent or sequential
    passes) gives the same chain state, samples and G(t) under every option setting the
    autotuner may choose.
"""
import numpy as np
import pytest

from test_gpu_parity import _bits, assert_same_state

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("case", [
    dict(n=300, nk=25, beta=1.0, cd=0, events=70, jumps=7),
    dict(n=3200, nk=40, beta=2.0, cd=1, events=40, jumps=16, exp_wait=True),
    dict(n=128, nk=30, beta=1.0, cd=1, events=50, jumps=4, kappa=[0, 0.05, 0, 0, 0, 0, 0, 0, 0]),
    dict(n=64, nk=300, beta=1.0, cd=1, events=60, jumps=9),
], ids=lambda c: "n%d_nk%d_j%d" % (c["n"], c["nk"], c["jumps"]))
def test_plain_multi_jump_parity(oracle, case):
    import gpu_dsm_amd as g
    kappa = case.get("kappa")
    with g.Engine(0) as e:
        e.set_option("plain_jumps", case["jumps"])
        assert e.get_option("plain_jumps") == case["jumps"]
        ens = e.add_ensemble(case["n"], case["nk"], case["beta"], case["cd"], kappa=kappa, seed=11, salt=1,
                             exp_wait=case.get("exp_wait", False))
        e.init_equilibrium(ens)
        P = oracle.Params(case["beta"], case["nk"], case["cd"], kappa=kappa, exp_wait=case.get("exp_wait", False))
        O = oracle.Ensemble(P, seed=11, n=case["n"], salt=1)
        st = e.advance_events(ens, case["events"], mode="plain")
        ev, _ = O.advance(max_events=case["events"])
        assert st["events"] == ev == case["n"] * case["events"]
        assert st["kmc_launches"] < case["events"]
        assert_same_state(e.state(ens), O.state(), "plain, %d jumps per launch" % case["jumps"])
        st = e.advance(ens, float(O.state()["t"].max()) + 3.0, mode="plain")
        O.advance(t_end=float(O.state()["t"].max()) + 3.0)
        assert_same_state(e.state(ens), O.state(), "plain to a time target")


SETTINGS = [
    {},
    dict(concurrent=0),
    dict(concurrent=1, split_margin=0),
    dict(concurrent=1, wide_pair=0),
    dict(concurrent=1, split_margin=4, blocks_per_cu=2),
    dict(small_per_cu=0, lds_table_floats=0),
    dict(small_per_cu=100000),
]


def test_options_do_not_change_results():
    import gpu_dsm_amd as g
    rng = np.random.default_rng(3)
    nk = rng.choice([20, 50, 90, 180, 400, 700], size=5000, p=[0.3, 0.3, 0.2, 0.12, 0.06, 0.02]).astype(np.int32)
    T, p, m = 12.0, 16, 2
    ref = None
    for opt in SETTINGS:
        with g.Engine(0) as e:
            for k, v in opt.items():
                e.set_option(k, v)
            ens = e.add_ensemble(nk.size, int(nk.max()), 1.0, 1, seed=8, sample="gt", delta=1.0, corr_p=p, corr_m=m,
                                 corr_levels=2, window=4, nk_per_chain=nk, partner="melt")
            e.init_equilibrium(ens)
            st = e.advance(ens, T)
            out = e.state(ens)
            out["events"] = st["events"]
            out["lag"], out["G"] = e.correlator(ens)
        if ref is None:
            ref = out
            continue
        assert out["events"] == ref["events"], opt
        assert_same_state(out, ref, "options %s" % opt)
        assert np.array_equal(_bits(out["G"]), _bits(ref["G"])), opt
