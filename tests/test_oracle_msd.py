"""Pins for the oracle's MSD(t) mode (CPU only).

MSD follows the centre of mass of each chain in the fixed frame: R1, the position of
the first entanglement, is carried through the events that create or destroy the
first entanglement (common/eq_ensemble_kernel.cu:447-455,531-539,606-614,645-653), a
sample is sum_i (P_i + Q_i/2) N_i/Nk + R1 with P_i the left end of strand i relative
to R1 (:297-323), and the multi-tau correlator carries the first value of every
block up a level and accumulates squared differences (:181-189,203,218;
common/correlator.cu:174-177).
"""
import numpy as np
import pytest


def _positions(st, i):
    """Fixed-frame positions of the entanglements of chain i (float64): entanglement e
    sits between strands e and e+1, X_0 = R1, X_e = R1 + Q_1 + ... + Q_e."""
    Z = int(st["Z"][i])
    Q = st["Q"][i, :Z].astype(np.float64)
    X = st["R1"][i].astype(np.float64) + np.concatenate([np.zeros((1, 3)), np.cumsum(Q[1:Z - 1], axis=0)])
    return X[:max(Z - 1, 0)]


def _one_removed_or_added(a, b, tol):
    """True when b equals a with one row removed or one row inserted (within tol)."""
    if abs(len(a) - len(b)) != 1:
        return False
    lo, hi = (a, b) if len(a) < len(b) else (b, a)
    for skip in range(len(hi)):
        if np.allclose(np.delete(hi, skip, axis=0), lo, rtol=0, atol=tol):
            return True
    return False


@pytest.mark.parametrize("nk,beta,cd", [(10, 1.0, 1), (16, 0.5, 1), (12, 2.0, 0)])
def test_entanglement_positions_survive_every_event(oracle, nk, beta, cd):
    """No event moves an entanglement that it does not create or destroy: after each
    jump the fixed-frame positions of the surviving entanglements are unchanged, and a
    created or destroyed one is the only difference.  A sign error, a wrong strand or a
    missing event kind in the R1 bookkeeping breaks this at the first head event."""
    P = oracle.Params(beta, nk, cd)
    E = oracle.Ensemble(P, seed=11, n=48)
    prev = E.state()
    changed = kept = 0
    for step in range(400):
        E.advance(max_events=1, msd=True)
        st = E.state()
        for i in range(E.n):
            a, b = _positions(prev, i), _positions(st, i)
            tol = 1e-4 * (1.0 + max(np.abs(a).max(initial=0.0), np.abs(b).max(initial=0.0)))
            if len(a) == len(b):
                assert np.allclose(a, b, rtol=0, atol=tol), (step, i)
                kept += 1
            else:
                assert _one_removed_or_added(a, b, tol), (step, i, len(a), len(b))
                changed += 1
        prev = st
    assert changed > 500 and kept > 1000


def test_com_sample_is_the_bead_average(oracle):
    """The sample equals the mean position of the chain's Nk Kuhn beads when strand i
    carries N_i beads evenly spaced along its vector from its left entanglement (the
    dangling strands have zero vectors), computed here bead by bead in float64."""
    P = oracle.Params(1.0, 30, 1)
    E = oracle.Ensemble(P, seed=5, n=40)
    E.advance(t_end=40.0, msd=True)
    st = E.state()
    # sample 0 of the current state (recorded before any further jump)
    E.set_next_sample(0)
    _, s = E.advance(max_events=0, delta=1.0, k0=0, k_end=1, samples=True, msd=True)
    k = 0
    assert np.all(st["Z"] > 2) and np.abs(st["R1"]).max() > 0
    for i in range(E.n):
        Z = int(st["Z"][i])
        Q = st["Q"][i, :Z].astype(np.float64)
        N = st["N"][i, :Z]
        start = st["R1"][i].astype(np.float64)
        beads = []
        for j in range(Z):
            for m in range(N[j]):
                beads.append(start + (m + 0.5) / N[j] * Q[j])
            start = start + Q[j]
        com = np.mean(beads, axis=0)
        np.testing.assert_allclose(s[k, i], com, rtol=2e-5, atol=2e-5 * (1 + np.abs(com).max()))


def _brute_msd(x, p, m, s):
    """Multi-tau MSD by definition: level k holds every m^k-th value of the series (the
    first of each block); point (k, j) averages |y_i - y_(i-j)|^2 over all available
    pairs, summed over the components."""
    out = {}
    for k in range(s):
        y = x[::m ** k]
        nb = len(x) // m ** k
        y = y[:nb]
        jmin = 0 if k == 0 else p // m
        for j in range(jmin, p):
            if j < len(y):
                d = y[j:] - y[:len(y) - j]
                out[(k, j)] = float(np.mean(np.sum(d * d, axis=-1)))
    return out


def test_msd_correlator_matches_definition(oracle):
    rng = np.random.default_rng(3)
    p, m, s, n, T = 8, 2, 5, 3, 200
    t = np.arange(T, dtype=np.float64)
    x = np.stack([0.01 * t[:, None] ** 2 + rng.normal(size=(T, 3)).cumsum(axis=0) for _ in range(n)], axis=1)
    cr = oracle.Correlator(p, m, s, n, type=1)
    for row in x:
        cr.add_row(row)
    lag, g = cr.result()
    ref = {}
    for c in range(n):
        for key, v in _brute_msd(x[:, c], p, m, s).items():
            ref[key] = ref.get(key, 0.0) + v / n
    im = 0
    for k in range(s):
        for j in range(0 if k == 0 else p // m, p):
            assert lag[im] == j * m ** k
            if (k, j) in ref:
                assert g[im] == pytest.approx(ref[(k, j)], rel=1e-12, abs=1e-12), (k, j)
            im += 1


def test_msd_grows_linearly_at_long_times(oracle):
    """Free diffusion of the centre of mass: for lags well past the disentanglement
    time of a short chain, MSD(t) grows linearly, MSD(2t)/MSD(t) = 2 within noise."""
    nk, n, T = 8, 512, 3000.0
    P = oracle.Params(1.0, nk, 1)
    E = oracle.Ensemble(P, seed=2, n=n)
    _, smp = E.advance(t_end=T, delta=1.0, k0=0, k_end=int(T), samples=True, msd=True)
    p, m = 16, 2
    s = oracle.corr_levels(T, 1.0, p, m)
    cr = oracle.Correlator(p, m, s, n, type=1)
    for row in smp:
        cr.add_row(row)
    lag, g = cr.result()
    gi = dict(zip(lag.tolist(), g.tolist()))
    assert gi[0.0] == 0.0
    r = gi[512.0] / gi[256.0]
    assert 1.7 < r < 2.3, r
