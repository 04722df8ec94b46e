"""Captured window passes (DSM_OPT_GRAPH): one window's launches -- parameter staging,
sample fills, the KMC pass (sequential or concurrent geometries), correlator update or
flow reduction -- captured as a HIP graph and replayed for later windows with their
parameters rewritten in a pinned block.  Every replay must give exactly the eager
result: each advance call below is one window (one replay, with the program's own host
work and synchronisation between replays), and the state, samples and G(t) or flow rows
are compared bit for bit with an engine that runs the same windows eagerly."""
import numpy as np
import pytest

from test_gpu_parity import _bits, assert_same_state

pytestmark = pytest.mark.gpu

CASES = {
    "gt_small_pass": [dict(n_chains=1000, nk=25, beta=1.0, cd=0, seed=1, sample="gt", delta=1.0, corr_p=16,
                           corr_m=2, corr_levels=4, window=8)],
    "gt_row_form": [dict(n_chains=8192, nk=60, beta=1.0, cd=1, seed=2, sample="gt", delta=1.0, corr_p=16, corr_m=2,
                         corr_levels=4, window=8)],
    "flow_four_ensembles": [dict(n_chains=n, nk=24, beta=1.0, cd=1, kappa=[0, r, 0, 0, 0, 0, 0, 0, 0], seed=5,
                                 salt=q, sample="flow", delta=1.0, window=8)
                            for q, (n, r) in enumerate([(1500, 0.0), (2560, 1e-3), (200, 1e-2), (720, 1e-1)])],
    "msd_polydisperse_concurrent": [dict(n_chains=3000, nk=600, beta=1.0, cd=1, seed=9, sample="msd", delta=1.0,
                                         corr_p=16, corr_m=2, corr_levels=4, window=8,
                                         nk_per_chain=np.random.default_rng(4).choice(
                                             [20, 80, 150, 300, 600], size=3000, p=[0.3, 0.3, 0.2, 0.15, 0.05]),
                                         partner="melt")],
}


def _run(descs, graph, windows, dt):
    import gpu_dsm_amd as g
    snaps = []
    with g.Engine(0) as e:
        e.set_option("graph", graph)
        ids = []
        for d in descs:
            ids.append(e.add_ensemble(**d))
            e.init_equilibrium(ids[-1])
        t = 0.0
        for w in range(windows):
            t += dt
            st = e.advance_many(ids, t)
            snap = dict(launches=st["kmc_launches"], events=st["events"])
            for q, ens in enumerate(ids):
                s = e.state(ens)
                if descs[q]["sample"] == "flow":
                    s["rows"] = e.flow_stress(ens)
                else:
                    s["lag"], s["G"] = e.correlator(ens)
                    s["k0"], s["last"] = e.last_samples(ens)
                snap[q] = s
            snaps.append(snap)
    return snaps


@pytest.mark.parametrize("name", list(CASES))
def test_graph_replays_match_eager(name):
    descs = CASES[name]
    windows, dt = 5, 8.0
    eager = _run(descs, 0, windows, dt)
    graph = _run(descs, 1, windows, dt)
    for w in range(windows):
        assert graph[w]["events"] == eager[w]["events"], "window %d" % w
        # a replay is one launch; the eager window launches every kernel of its pass
        assert graph[w]["launches"] == 1 and eager[w]["launches"] >= 1
        for q in range(len(descs)):
            a, b = graph[w][q], eager[w][q]
            assert_same_state(a, b, "%s window %d ensemble %d" % (name, w, q))
            if "rows" in b:
                assert np.array_equal(a["rows"], b["rows"])
            else:
                assert np.array_equal(_bits(a["last"]), _bits(b["last"]))
                assert np.array_equal(a["G"].view(np.uint64), b["G"].view(np.uint64))
