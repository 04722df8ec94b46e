"""HIP path vs the CPU oracle (mirror-FP32 mode), through the C ABI.

Integer chain state (Z, every N_i, event counts) and the FP32 strand vectors,
CD rates and FP64 clocks must be bit-identical after a fixed number of jumps.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


def _bits(a):
    return np.ascontiguousarray(a, np.float32).view(np.uint32)


def assert_same_state(g, o, what=""):
    Z = g["Z"]
    assert np.array_equal(Z, o["Z"]), "%s: Z differs in %d chains (first %s)" % (
        what, int((Z != o["Z"]).sum()), np.nonzero(Z != o["Z"])[0][:5])
    cap = min(g["N"].shape[1], o["N"].shape[1])
    k = np.arange(cap)[None, :]
    on = k < Z[:, None]
    link = k < (Z[:, None] - 1)
    gN, oN = np.where(on, g["N"][:, :cap], 0), np.where(on, o["N"][:, :cap], 0)
    bad = np.nonzero((gN != oN).any(axis=1))[0]
    assert bad.size == 0, "%s: N differs in chains %s" % (what, bad[:5])
    gQ = np.where(on[..., None], _bits(g["Q"][:, :cap]), 0)
    oQ = np.where(on[..., None], _bits(o["Q"][:, :cap]), 0)
    bad = np.nonzero((gQ != oQ).any(axis=(1, 2)))[0]
    assert bad.size == 0, "%s: Q bits differ in chains %s" % (what, bad[:5])
    gw = np.where(link, _bits(g["w"][:, :cap]), 0)
    ow = np.where(link, _bits(o["w"][:, :cap]), 0)
    bad = np.nonzero((gw != ow).any(axis=1))[0]
    assert bad.size == 0, "%s: CD rates differ in chains %s" % (what, bad[:5])
    assert np.array_equal(g["t"], o["t"]), "%s: FP64 clocks differ (max |d| %g)" % (
        what, float(np.abs(g["t"] - o["t"]).max()))
    assert np.array_equal(g["ev"], o["ev"]), "%s: event counters differ" % what


# ------------------------------------------------------------------ building blocks
def test_philox_matches_rocrand_vectors():
    import gpu_dsm_amd as g
    vec = json.load(open(os.path.join(GOLDEN, "philox_rocrand.json")))
    ck = np.array([v["ctr"] + v["key"] for v in vec], np.uint32)
    out = g.probe_philox(ck)
    assert np.array_equal(out, np.array([v["out"] for v in vec], np.uint32))


def test_scan_tree_matches_oracle(oracle):
    """Every chain geometry of the KMC kernels sums in the oracle's canonical tree."""
    import gpu_dsm_amd as g
    rng = np.random.default_rng(5)
    for n in (1, 3, 17, 64, 65, 200, 256, 257, 640, 1000, 1024):
        for vals in (rng.random(n).astype(np.float32),
                     (rng.standard_normal(n) * 10).astype(np.float32),
                     rng.exponential(size=n).astype(np.float32) * np.float32(1e3)):
            gpu, tot = g.probe_scan(vals, form=1)
            assert np.array_equal(_bits(gpu), _bits(oracle.scan_tree(vals)))
            assert _bits(np.float32(tot)) == _bits(np.float32(oracle.tree_total(vals)))
            if n <= 512:
                # the wave form with two register groups (its own carry code)
                gpu, tot = g.probe_scan(vals, form=5)
                assert np.array_equal(_bits(gpu), _bits(oracle.scan_tree(vals)))
                assert _bits(np.float32(tot)) == _bits(np.float32(oracle.tree_total(vals)))
    for n in (2, 100, 128, 129, 191, 192, 193, 255, 320, 384, 385, 449, 511, 512):
        vals = (rng.exponential(size=n) * 20).astype(np.float32)
        gpu, tot = g.probe_scan(vals, form=5)
        assert np.array_equal(_bits(gpu), _bits(oracle.scan_tree(vals)))
        assert _bits(np.float32(tot)) == _bits(np.float32(oracle.tree_total(vals)))
    for n in (1, 5, 40, 63, 64):
        vals = (rng.standard_normal((4, n)) * 3).astype(np.float32)
        gpu, tot = g.probe_scan(vals, form=0)
        for r in range(4):
            assert np.array_equal(_bits(gpu[r]), _bits(oracle.scan_tree(vals[r])))
            assert _bits(np.float32(tot[r])) == _bits(np.float32(oracle.tree_total(vals[r])))
    for n in (1, 7, 63, 64, 65, 100, 127, 128):
        vals = (rng.exponential(size=(2, n)) * 50).astype(np.float32)
        gpu, tot = g.probe_scan(vals, form=2)
        for r in range(2):
            assert np.array_equal(_bits(gpu[r]), _bits(oracle.scan_tree(vals[r])))
            assert _bits(np.float32(tot[r])) == _bits(np.float32(oracle.tree_total(vals[r])))
    # pair form with two register groups: block carries across rows and groups
    for n in (1, 64, 65, 128, 129, 130, 191, 192, 193, 200, 255, 256):
        vals = (rng.exponential(size=(2, n)) * 50).astype(np.float32)
        gpu, tot = g.probe_scan(vals, form=3)
        for r in range(2):
            assert np.array_equal(_bits(gpu[r]), _bits(oracle.scan_tree(vals[r])))
            assert _bits(np.float32(tot[r])) == _bits(np.float32(oracle.tree_total(vals[r])))
    # row form with two register groups: the block carry across the two groups of a row
    for n in (1, 5, 63, 64, 65, 66, 100, 127, 128):
        vals = (rng.exponential(size=(4, n)) * 50).astype(np.float32)
        gpu, tot = g.probe_scan(vals, form=4)
        for r in range(4):
            assert np.array_equal(_bits(gpu[r]), _bits(oracle.scan_tree(vals[r])))
            assert _bits(np.float32(tot[r])) == _bits(np.float32(oracle.tree_total(vals[r])))


@pytest.mark.parametrize("which", ["exp", "log", "sin2pi", "cos2pi"])
def test_math_bit_exact(oracle, which):
    import gpu_dsm_amd as g
    rng = np.random.default_rng(11)
    if which == "exp":
        x = np.concatenate([np.linspace(-90, 90, 200001), rng.standard_normal(100000) * 3,
                            [0.0, -0.0, 88.0, -87.0, 1e-30, -1e-30]]).astype(np.float32)
    elif which == "log":
        x = np.concatenate([((np.arange(0, 1 << 24, 97) + 0.5) * 2.0 ** -24),
                            np.exp(rng.uniform(-80, 80, 100000)), [1.0, 2.0, 0.5, 3.4e38]]).astype(np.float32)
    else:
        x = ((np.arange(0, 1 << 24, 53) + 0.5) * 2.0 ** -24).astype(np.float32)
    assert np.array_equal(_bits(g.probe_math(x, which)), _bits(oracle.vec_math(which, x)))


# ------------------------------------------------------------------ equilibrium init
INIT_CASES = [
    dict(n=500, nk=25, beta=1.0, cd=0),
    dict(n=300, nk=100, beta=1.0, cd=1),
    dict(n=200, nk=60, beta=0.5, cd=1),
    dict(n=64, nk=300, beta=2.0, cd=1),
]


@pytest.mark.parametrize("case", INIT_CASES, ids=lambda c: "nk%d_b%g_cd%d" % (c["nk"], c["beta"], c["cd"]))
def test_init_parity(engine, oracle, case):
    P = oracle.Params(case["beta"], case["nk"], case["cd"])
    ens = engine.add_ensemble(case["n"], case["nk"], case["beta"], case["cd"], seed=7, salt=3)
    engine.init_equilibrium(ens)
    O = oracle.Ensemble(P, seed=7, n=case["n"], salt=3)
    assert_same_state(engine.state(ens), O.state(), "init")
    engine.remove_ensemble(ens)


# ------------------------------------------------------------------ KMC jumps
KMC_CASES = [
    dict(n=1, nk=30, beta=1.0, cd=1, events=200),      # a single chain
    dict(n=300, nk=25, beta=1.0, cd=0, events=400),
    dict(n=200, nk=100, beta=1.0, cd=1, events=300),
    dict(n=127, nk=50, beta=0.5, cd=1, events=300),
    dict(n=96, nk=160, beta=1.0, cd=1, events=250),
    dict(n=64, nk=240, beta=1.0, cd=1, events=300),  # Z ~ 120: pair form (Z < 128) and wave form, crossings
    dict(n=48, nk=300, beta=1.0, cd=1, events=200),
    dict(n=9, nk=1000, beta=0.5, cd=1, events=60),
    dict(n=5, nk=2000, beta=0.5, cd=1, events=30),   # Z ~ 1330: the Z <= 2048 wave geometry
    # thousands of work units: several per persistent wave, and chains that reach the
    # row form's capacity (Z = 64) mid-pass and continue in the wave form
    dict(n=4096, nk=100, beta=1.0, cd=1, events=80),
    dict(n=16384, nk=40, beta=1.0, cd=1, events=40),
    dict(n=128, nk=40, beta=2.0, cd=1, events=300, exp_wait=True),
    dict(n=128, nk=30, beta=1.0, cd=1, events=300, kappa=[0, 0.05, 0, 0, 0, 0, 0, 0, 0]),
    dict(n=64, nk=20, beta=1.0, cd=1, events=300, kappa=[0.02, 0, 0, 0, -0.01, 0, 0, 0, -0.01]),
    # more chains than one wave-form wave each (12 per CU): batched passes use the row form
    dict(n=3200, nk=40, beta=2.0, cd=1, events=60, exp_wait=True),
    dict(n=3200, nk=30, beta=1.0, cd=1, events=60, kappa=[0, 0.05, 0, 0, 0, 0, 0, 0, 0]),
    dict(n=3200, nk=20, beta=1.0, cd=1, events=60, kappa=[0.02, 0, 0, 0, -0.01, 0, 0, 0, -0.01]),
]


def _case_id(c):
    s = "n%d_nk%d_b%g_cd%d" % (c["n"], c["nk"], c["beta"], c["cd"])
    if c.get("exp_wait"):
        s += "_expwait"
    if c.get("kappa"):
        s += "_flow"
    return s


@pytest.mark.parametrize("mode", ["batched", "plain"])
@pytest.mark.parametrize("case", KMC_CASES, ids=_case_id)
def test_kmc_parity(engine, oracle, case, mode):
    kappa = case.get("kappa")
    P = oracle.Params(case["beta"], case["nk"], case["cd"], kappa=kappa, exp_wait=case.get("exp_wait", False))
    ens = engine.add_ensemble(case["n"], case["nk"], case["beta"], case["cd"], kappa=kappa, seed=11, salt=1,
                              exp_wait=case.get("exp_wait", False))
    engine.init_equilibrium(ens)
    O = oracle.Ensemble(P, seed=11, n=case["n"], salt=1)
    half = case["events"] // 2
    for k, n_ev in enumerate((half, case["events"] - half)):
        st = engine.advance_events(ens, n_ev, mode=mode)
        ev, _ = O.advance(max_events=n_ev)
        assert st["events"] == ev, "event totals differ (%d vs %d)" % (st["events"], ev)
        assert_same_state(engine.state(ens), O.state(), "%s after chunk %d" % (mode, k))
    engine.remove_ensemble(ens)


GEX = (1.5, 1.0, 40000.0, 120.0)  # a, b, mp, Mk


@pytest.mark.parametrize("n", [600, 3200])
@pytest.mark.parametrize("mode", ["batched", "plain"])
@pytest.mark.parametrize("partner", ["gex", "melt"])
def test_partner_background_parity(engine, oracle, partner, mode, n):
    """Polydisperse CD background: partner tables identical to the oracle's, and every
    CD lifetime drawn through them bit-exact (initial links and created ones), on a
    polydisperse ensemble whose chains use both the row and the wave form."""
    rng = np.random.default_rng(7)
    nk = rng.choice([12, 30, 60, 140, 400], size=n, p=[0.3, 0.3, 0.2, 0.15, 0.05]).astype(np.int32)
    nmax = int(nk.max())
    P = oracle.Params(1.0, nmax, 1, nmax=nmax, partner=partner, gex=GEX, nk_chains=nk)
    ens = engine.add_ensemble(n, nmax, 1.0, 1, seed=13, salt=2, nk_per_chain=nk, partner=partner, gex=GEX)
    for name in ("PF", "PD", "PF_NK", "PD_NK", "CC"):
        assert np.array_equal(_bits(engine.table(ens, name)), _bits(P.table(name))), name
    engine.init_equilibrium(ens)
    O = oracle.Ensemble(P, seed=13, n=n, salt=2, nk_per_chain=nk)
    assert_same_state(engine.state(ens), O.state(), "initial")
    for k, n_ev in enumerate((60, 90)):
        st = engine.advance_events(ens, n_ev, mode=mode)
        ev, _ = O.advance(max_events=n_ev)
        assert st["events"] == ev
        assert_same_state(engine.state(ens), O.state(), "%s %s chunk %d" % (partner, mode, k))
    engine.remove_ensemble(ens)


def test_batched_time_target_parity(engine, oracle):
    """Run to a time target (persistent form with Z-binning) and compare."""
    P = oracle.Params(1.0, 100, 1)
    ens = engine.add_ensemble(256, 100, 1.0, 1, seed=3)
    engine.init_equilibrium(ens)
    O = oracle.Ensemble(P, seed=3, n=256)
    for t in (5.0, 12.5, 30.0):
        engine.advance(ens, t)
        O.advance(t_end=t)
        assert_same_state(engine.state(ens), O.state(), "t=%g" % t)
    engine.remove_ensemble(ens)


# ------------------------------------------------------------------ samples, G(t), flow stress
def test_gt_samples_and_correlator(engine, oracle):
    n, nk, T = 200, 25, 300.0
    p, m = 16, 2
    s = oracle.corr_levels(T, 1.0, p, m)
    P = oracle.Params(1.0, nk, 0)
    ens = engine.add_ensemble(n, nk, 1.0, 0, seed=1, sample="gt", delta=1.0, corr_p=p, corr_m=m,
                              corr_levels=s, window=64)
    engine.init_equilibrium(ens)
    O = oracle.Ensemble(P, seed=1, n=n)
    _, samples = O.advance(t_end=T, delta=1.0, k0=0, k_end=int(T), samples=True)
    engine.advance(ens, T)
    assert_same_state(engine.state(ens), O.state(), "G(t) run")
    k0, last = engine.last_samples(ens)
    assert np.array_equal(_bits(last), _bits(samples[k0:k0 + last.shape[0]]))
    corr = oracle.Correlator(p, m, s, n)
    for row in samples:
        corr.add_row(row)
    lag_o, g_o = corr.result()
    lag, g = engine.correlator(ens)
    assert np.array_equal(lag, lag_o)
    np.testing.assert_allclose(g, g_o, rtol=1e-10, atol=1e-12)
    engine.remove_ensemble(ens)


@pytest.mark.parametrize("Z,geom", [(1, 0), (2, 0), (60, 0), (64, 0), (1, 1), (63, 1), (65, 1), (100, 1),
                                     (128, 1), (65, 2), (250, 2), (256, 2), (257, 3), (268, 3), (500, 3),
                                     (512, 3), (513, 4), (900, 4), (1024, 4), (1, 7), (100, 7), (128, 7),
                                     (129, 7), (130, 7), (200, 7), (256, 7), (1, 8), (63, 8), (64, 8),
                                     (65, 8), (100, 8), (128, 8)])
def test_com_sample_matches_oracle(oracle, Z, geom):
    """The centre-of-mass sample of the kernels, in every geometry, against the oracle."""
    import gpu_dsm_amd as g
    rng = np.random.default_rng(Z)
    N = rng.integers(1, 4, Z)
    Q = rng.normal(size=(Z, 3)).astype(np.float32)
    if Z > 1:
        Q[0] = 0.0
        Q[Z - 1] = 0.0
    nk = int(N.sum())
    E = oracle.Ensemble(oracle.Params(1.0, nk, 0), seed=1, n=1)
    cap = E.cap
    Nf = np.zeros((1, cap), np.int32); Nf[0, :Z] = N
    Qf = np.zeros((1, cap, 3), np.float32); Qf[0, :Z] = Q
    E.set_state(np.array([Z]), Nf, Qf, np.zeros((1, cap), np.float32), t=np.array([0.5]),
                next_sample=np.zeros(1, np.int64))
    _, smp = E.advance(max_events=0, delta=1.0, k0=0, k_end=1, samples=True, msd=True)
    got = g.probe_com(Q, N, geom)
    assert np.array_equal(_bits(got), _bits(smp[0, 0])), (got, smp[0, 0])


MSD_CASES = [
    dict(n=200, nk=12, beta=1.0, cd=1, T=300.0, mode="batched"),   # short chains: head events every few jumps
    dict(n=127, nk=40, beta=0.5, cd=1, T=150.0, mode="batched"),
    dict(n=64, nk=12, beta=1.0, cd=0, T=40.0, mode="plain"),
    dict(n=9, nk=400, beta=0.5, cd=1, T=15.0, mode="batched"),     # wave form (Z up to ~290)
    dict(n=3200, nk=40, beta=0.5, cd=1, T=20.0, mode="batched"),   # row form (a pass of many chains)
]


@pytest.mark.parametrize("case", MSD_CASES,
                         ids=lambda c: "n%d_nk%d_b%g_cd%d_%s" % (c["n"], c["nk"], c["beta"], c["cd"], c["mode"]))
def test_msd_samples_and_correlator(engine, oracle, case):
    """MSD(t): R1 bookkeeping, centre-of-mass samples (bit-exact) and the type-1 multi-tau
    correlator against the oracle."""
    n, nk, beta, cd, T = case["n"], case["nk"], case["beta"], case["cd"], case["T"]
    p, m = 16, 2
    s = oracle.corr_levels(T, 1.0, p, m)
    P = oracle.Params(beta, nk, cd)
    ens = engine.add_ensemble(n, nk, beta, cd, seed=8, sample="msd", delta=1.0, corr_p=p, corr_m=m,
                              corr_levels=s, window=64)
    engine.init_equilibrium(ens)
    O = oracle.Ensemble(P, seed=8, n=n)
    _, samples = O.advance(t_end=T, delta=1.0, k0=0, k_end=int(T), samples=True, msd=True)
    engine.advance(ens, T, mode=case["mode"])
    g_st, o_st = engine.state(ens), O.state()
    assert_same_state(g_st, o_st, "MSD run")
    assert np.array_equal(_bits(g_st["R1"]), _bits(o_st["R1"])), "R1 differs"
    assert np.abs(o_st["R1"]).max() > 0
    k0, last = engine.last_samples(ens)
    ref = samples[k0:k0 + last.shape[0]]
    bad = np.argwhere(_bits(last) != _bits(ref))
    assert bad.size == 0, "%d sample values differ; first (row, chain, comp) %s, gpu %s oracle %s, Z %s" % (
        bad.shape[0], bad[:4].tolist(), [float(last[tuple(b)]) for b in bad[:4]],
        [float(ref[tuple(b)]) for b in bad[:4]], g_st["Z"][np.unique(bad[:, 1])][:6].tolist())
    corr = oracle.Correlator(p, m, s, n, type=1)
    for row in samples:
        corr.add_row(row)
    lag_o, g_o = corr.result()
    lag, g = engine.correlator(ens)
    assert np.array_equal(lag, lag_o)
    np.testing.assert_allclose(g, g_o, rtol=1e-10, atol=1e-12)
    assert g[0] == 0.0 and g[1] > 0
    engine.remove_ensemble(ens)


def test_flow_stress_rows(engine, oracle):
    n, nk, T, dlt = 128, 30, 20.0, 2.0
    kappa = [0, 0.02, 0, 0, 0, 0, 0, 0, 0]
    P = oracle.Params(1.0, nk, 1, kappa=kappa)
    ens = engine.add_ensemble(n, nk, 1.0, 1, kappa=kappa, seed=5, sample="flow", delta=dlt, window=4)
    engine.init_equilibrium(ens)
    O = oracle.Ensemble(P, seed=5, n=n)
    k_end = int(T / dlt) + 1
    _, samples = O.advance(t_end=T, delta=dlt, k0=1, k_end=k_end, samples=True, flow_samples=True)
    engine.advance(ens, T)
    assert_same_state(engine.state(ens), O.state(), "flow run")
    rows = engine.flow_stress(ens)
    assert rows.shape[0] == k_end - 1
    ref = samples.astype(np.float64).sum(axis=1)
    np.testing.assert_allclose(rows[:, :8], ref, rtol=1e-12, atol=1e-9)
    assert np.all(rows[:, 8] == n)
    engine.remove_ensemble(ens)


@pytest.mark.parametrize("scale", [1, 40])
@pytest.mark.parametrize("mode", ["batched", "plain"])
def test_multi_ensemble_flow_batch(engine, oracle, mode, scale):
    """Several ensembles (different shear rates, chain counts that are not multiples of
    four, one equilibrium ensemble) advanced together in one pass: each must match its
    own oracle run bit for bit, with its own flow output rows."""
    nk, T, dlt = 24, 12.0, 3.0
    rates = [0.0, 0.01, 0.1, 0.5]
    counts = [37 * scale, 64 * scale, 5 * scale, 18 * scale]  # scale 40: row form in batched passes
    ids, refs = [], []
    for q, (gd, n) in enumerate(zip(rates, counts)):
        kappa = [0, gd, 0, 0, 0, 0, 0, 0, 0]
        P = oracle.Params(1.0, nk, 1, kappa=kappa)
        ens = engine.add_ensemble(n, nk, 1.0, 1, kappa=kappa, seed=40 + q, salt=q, sample="flow", delta=dlt,
                                  window=3)
        engine.init_equilibrium(ens)
        O = oracle.Ensemble(P, seed=40 + q, n=n, salt=q)
        k_end = int(T / dlt) + 1
        _, samples = O.advance(t_end=T, delta=dlt, k0=1, k_end=k_end, samples=True, flow_samples=True)
        ids.append(ens)
        refs.append((O, samples, n, k_end))
    engine.advance_many(ids, T, mode=mode)
    for ens, (O, samples, n, k_end) in zip(ids, refs):
        assert_same_state(engine.state(ens), O.state(), "ensemble %d" % ens)
        rows = engine.flow_stress(ens)
        assert rows.shape[0] == k_end - 1
        np.testing.assert_allclose(rows[:, :8], samples.astype(np.float64).sum(axis=1), rtol=1e-12, atol=1e-9)
    for ens in ids:
        engine.remove_ensemble(ens)


def test_many_small_ensembles_batch(engine, oracle):
    """Many small ensembles of different lengths in one batched pass (single chains,
    counts that are not multiples of four, chains in the row and the wave form): a
    16-lane segment that takes a chain of another ensemble holds it until its wave has
    switched ensembles, so every chain must still match its own oracle run bit for bit."""
    counts = [1, 2, 1, 7, 33, 1, 130, 3, 64, 1]
    nks = [20, 90, 150, 50, 100, 8, 70, 300, 40, 60]
    T = 20.0
    ids, refs = [], []
    for q, (n, nk) in enumerate(zip(counts, nks)):
        P = oracle.Params(1.0, nk, 1)
        ens = engine.add_ensemble(n, nk, 1.0, 1, seed=60 + q, salt=q)
        engine.init_equilibrium(ens)
        O = oracle.Ensemble(P, seed=60 + q, n=n, salt=q)
        O.advance(t_end=T)
        ids.append(ens)
        refs.append(O)
    engine.advance_many(ids, T)
    for ens, O in zip(ids, refs):
        assert_same_state(engine.state(ens), O.state(), "ensemble %d" % ens)
    for ens in ids:
        engine.remove_ensemble(ens)


@pytest.mark.parametrize("sample", ["gt", "msd"])
def test_correlator_beside_wave_form_tail(engine, oracle, sample):
    """Windows in which chains outgrow the row form (Nk = 100, enough chains for the row
    form): the chains that reach Z = 64 record the rest of their window's samples in the
    wave form after the row form has finished.  Samples bit-exact and the correlator equal
    to the oracle's over the same samples."""
    n, nk, T, win = 4096, 100, 48.0, 16
    p, m = 16, 2
    s = oracle.corr_levels(T, 1.0, p, m)
    P = oracle.Params(1.0, nk, 1)
    ens = engine.add_ensemble(n, nk, 1.0, 1, seed=17, sample=sample, delta=1.0, corr_p=p, corr_m=m,
                              corr_levels=s, window=win)
    engine.init_equilibrium(ens)
    O = oracle.Ensemble(P, seed=17, n=n)
    _, samples = O.advance(t_end=T, delta=1.0, k0=0, k_end=int(T), samples=True, msd=(sample == "msd"))
    z_max = 0
    t = 0.0
    while t < T:
        t = min(T, t + win)
        engine.advance(ens, t)
        z_max = max(z_max, int(engine.state(ens)["Z"].max()))
    assert z_max >= 64, "no chain reached the wave form (max Z %d)" % z_max
    assert_same_state(engine.state(ens), O.state(), "%s run" % sample)
    k0, last = engine.last_samples(ens)
    assert np.array_equal(_bits(last), _bits(samples[k0:k0 + last.shape[0]]))
    corr = oracle.Correlator(p, m, s, n, type=1 if sample == "msd" else 0)
    for row in samples:
        corr.add_row(row)
    lag_o, g_o = corr.result()
    lag, g = engine.correlator(ens)
    assert np.array_equal(lag, lag_o)
    np.testing.assert_allclose(g, g_o, rtol=1e-10, atol=1e-12)
    engine.remove_ensemble(ens)
