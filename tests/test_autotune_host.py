"""The autotuner's workload signature and cache (host only)."""
import numpy as np


def test_signature_separates_what_the_layout_depends_on():
    from gpu_dsm_amd import autotune as A
    base = dict(n_chains=1000, nk=25, beta=1.0, cd=0, sample="gt")
    s0 = A.signature([base])
    assert s0 == A.signature([dict(base, seed=7)])                      # the seed does not matter
    assert s0 != A.signature([dict(base, nk=100)])                       # chain length
    assert s0 != A.signature([dict(base, n_chains=131072)])              # chain count
    assert s0 != A.signature([dict(base, kappa=[0, 1e-3, 0, 0, 0, 0, 0, 0, 0])])  # flow
    assert s0 != A.signature([dict(base, beta=2.0)])
    assert s0 != A.signature([dict(base, partner="melt", nk_per_chain=np.full(1000, 25))])
    poly = dict(base, nk=400, nk_per_chain=np.r_[np.full(500, 20), np.full(500, 400)])
    assert A.signature([poly]) != A.signature([dict(base, nk=210)])       # polydispersity
    assert A.signature([base] * 16) != s0                                 # ensembles per pass


def test_cache_roundtrip(tmp_path):
    from gpu_dsm_amd import autotune as A
    path = str(tmp_path / "sub" / "tune.json")
    A._store(path, "k1", dict(best=dict(A.DEFAULTS), default_rate=1.0, best_rate=1.0))
    A._store(path, "k2", dict(best=dict(A.DEFAULTS, window=64), default_rate=1.0, best_rate=1.2))
    d = A._load(path)
    assert set(d) == {"k1", "k2"} and d["k2"]["best"]["window"] == 64


def test_settle_keeps_defaults_at_their_better_rate():
    from gpu_dsm_amd import autotune as A
    other = dict(A.DEFAULTS, split_margin=0)
    # the sweep found nothing better: the defaults, at the better of their two timings
    best, rate, noise = A._settle(dict(A.DEFAULTS), 100.0, 100.0, 101.0, 0.03)
    assert best == A.DEFAULTS and rate == 101.0 and abs(noise - 0.01) < 1e-12
    # a setting within the gain threshold of the defaults does not replace them
    best, rate, _ = A._settle(other, 102.0, 100.0, 101.0, 0.03)
    assert best == A.DEFAULTS and rate == 101.0
    # a clear gain, above both the threshold and the run-to-run spread, does
    best, rate, _ = A._settle(other, 110.0, 100.0, 101.0, 0.03)
    assert best == other and rate == 110.0
    # ... unless the defaults' own two timings spread by more than that gain
    best, rate, _ = A._settle(other, 110.0, 100.0, 108.0, 0.03)
    assert best == A.DEFAULTS and rate == 108.0
