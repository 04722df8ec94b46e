"""Pins for the oracle's physics (CPU only): closed forms and invariants of the DSM.

References: p_CD fits (common/pcd_tau.h:32-82), rate formulas
(common/eq_ensemble_kernel.cu:97-126, :340-394), equilibrium init laws
(common/chain.cu:38-102), correlator (common/eq_ensemble_kernel.cu:133-227).
"""
import math

import numpy as np
import pytest
from scipy import special, stats


# ------------------------------------------------------------------ p_CD
def _pcd_double(beta, nk):
    """p_CD fit constants (common/pcd_tau.h:36-63) evaluated in double."""
    g = 0.667
    if beta != 1.0:
        z = (nk + beta) / (beta + 1.0)
        alpha = (0.053 * math.log(beta) + 0.31) * z ** (-0.012 * math.log(beta) - 0.024)
        t0 = 0.285 * (beta + 2) ** 0.515
        tmax = t0 if nk < 2 else 0.025 * (beta + 2) ** 2.6 * z ** 2.83
        td = t0 if nk < 2 else 0.036 * (beta + 2) ** 3.07 * (z - 1) ** 3.02
    else:
        alpha = 0.267096 - 0.375571 * math.exp(-0.0838237 * nk)
        t0 = 0.460277 + 0.298913 * math.exp(-0.0705314 * nk)
        tmax = t0 if nk < 4 else 0.0156137 * nk ** 3.18849
        td = t0 if nk < 4 else 0.0740131 * nk ** 3.18363
    return dict(g=g, alpha=alpha, tau_0=t0, tau_max=tmax, tau_d=td)


@pytest.mark.parametrize("beta,nk", [(1.0, 25), (1.0, 100), (1.0, 3), (0.5, 60), (2.0, 300), (1.0, 1000)])
def test_pcd_constants(oracle, beta, nk):
    P = oracle.Params(beta, nk, 1)
    s = P.scalars()
    ref = _pcd_double(beta, nk)
    for k in ("alpha", "tau_0", "tau_max", "tau_d"):
        assert s[k] == pytest.approx(ref[k], rel=3e-5), k
    a, t0, tm, td, g = ref["alpha"], ref["tau_0"], ref["tau_max"], ref["tau_d"], ref["g"]
    if tm > t0:
        # <1/tau> of p(tau) = (1-g) a t^(a-1)/(tm^a - t0^a) on [t0, tm] + g delta(t - td)
        from scipy.integrate import quad
        dens = lambda t: (1 - g) * a * t ** (a - 1) / (tm ** a - t0 ** a)
        lo, hi = math.log(t0), math.log(tm)
        # integrate in x = ln t so the power law is smooth over many decades
        mean_inv = quad(lambda x: dens(math.exp(x)), lo, hi, limit=400)[0] + g / td
        norm = quad(lambda x: dens(math.exp(x)) * math.exp(x), lo, hi, limit=400)[0] + g
        assert s["normdt"] == pytest.approx(mean_inv, rel=1e-4)
        assert norm == pytest.approx(1.0, abs=1e-7)
    # CD creation prefactor c = <1/tau>/beta (common/ensemble.cu:148)
    assert s["c_create"] == pytest.approx(s["normdt"] / beta, rel=1e-6)


@pytest.mark.parametrize("beta,nk", [(1.0, 100), (0.5, 60)])
def test_pcd_samplers(oracle, beta, nk):
    P = oracle.Params(beta, nk, 1)
    s = P.scalars()
    a, t0, tm, td, g = s["alpha"], s["tau_0"], s["tau_max"], s["tau_d"], s["g"]
    u = (np.arange(40000) + 0.5) / 40000.0
    tau_ft = 1.0 / np.array([P.rate_ft(x) for x in u])
    # f_t samples p(tau): CDF (1-g)(t^a - t0^a)/(tm^a - t0^a) below tm, atom g at td
    cont = tau_ft[tau_ft < 0.99 * td]
    assert len(cont) / len(u) == pytest.approx(1 - g, abs=1e-3)
    cdf = lambda t: np.clip((t ** a - t0 ** a) / (tm ** a - t0 ** a), 0, 1)
    assert stats.kstest(cont, cdf).statistic < 2e-3
    assert np.mean(1.0 / tau_ft) == pytest.approx(s["normdt"], rel=2e-3)
    # f_d_t samples p(tau)/(tau <1/tau>): atom at td of weight (g/td)/<1/tau>, continuous
    # part with density ~ t^(a-2) on [t0, tm]
    tau_fdt = 1.0 / np.array([P.rate_fdt(x) for x in u])
    atom = (g / td) / s["normdt"]
    assert (tau_fdt > 0.99 * td).mean() == pytest.approx(atom, abs=2.0 / len(u))
    cont = tau_fdt[tau_fdt < 0.99 * td]
    cdf2 = lambda t: np.clip((t ** (a - 1) - t0 ** (a - 1)) / (tm ** (a - 1) - t0 ** (a - 1)), 0, 1)
    assert stats.kstest(cont, cdf2).statistic < 2e-3


# ------------------------------------------------------------------ tables
def test_rate_tables_match_formulas(oracle):
    beta, nk = 0.7, 120
    P = oracle.Params(beta, nk, 1)
    N = np.arange(1, nk + 2, dtype=np.float64)
    chk = {
        "SP": 0.75 / (N * (N + 1)),
        "PP": (N / (N + 1)) ** 0.75,
        "EC": 2.0 / (beta * (N + 0.5)),
        "ED": 1.0 / (0.75 + 0.5 * N),
        "ED2": 1.0 / (N + 1.0),
        "T3": 3.0 / N,
    }
    for name, ref in chk.items():
        t = P.table(name)[1:nk + 2].astype(np.float64)
        np.testing.assert_allclose(t, ref, rtol=3e-7, err_msg=name)
    sm = P.table("SM")[2:nk + 2]
    np.testing.assert_allclose(sm, 0.75 / (N[1:] * (N[1:] - 1)), rtol=3e-7)
    pm = P.table("PM")[2:nk + 2]
    np.testing.assert_allclose(pm, (N[1:] / (N[1:] - 1)) ** 0.75, rtol=3e-7)
    fr = P.table("FR")
    x = np.arange(1, len(fr))
    np.testing.assert_allclose(fr[1:], 4.0 / x, rtol=1e-7)


# ------------------------------------------------------------------ detailed balance (App. C4)
def _gauss_weight(N, q2):
    return N ** -1.5 * math.exp(-1.5 * q2 / N)


def test_shift_detailed_balance(oracle):
    """W+(s) pi(s) == W-(s') pi(s') for s' = s with one Kuhn step moved j -> i."""
    P = oracle.Params(1.0, 200, 0)
    rng = np.random.default_rng(2)
    worst = 0.0
    for _ in range(4000):
        Ni, Nj = int(rng.integers(1, 40)), int(rng.integers(2, 40))
        endi, endj = bool(rng.random() < 0.3), bool(rng.random() < 0.3)
        q2i = 0.0 if endi else float(np.float32(rng.gamma(1.5, 2.0 * Ni / 3.0)))
        q2j = 0.0 if endj else float(np.float32(rng.gamma(1.5, 2.0 * Nj / 3.0)))
        wp, _ = P.pair_rates(Ni, Nj, q2i, q2j, endi, endj)
        _, wm = P.pair_rates(Ni + 1, Nj - 1, q2i, q2j, endi, endj)
        pi_s = (1.0 if endi else _gauss_weight(Ni, q2i)) * (1.0 if endj else _gauss_weight(Nj, q2j))
        pi_t = (1.0 if endi else _gauss_weight(Ni + 1, q2i)) * (1.0 if endj else _gauss_weight(Nj - 1, q2j))
        lhs, rhs = wp * pi_s, wm * pi_t
        worst = max(worst, abs(lhs - rhs) / max(lhs, rhs))
    assert worst < 2e-6


def test_end_moves_balance_is_one_over_beta(oracle):
    for beta in (0.5, 1.0, 2.0):
        nk = 80
        P = oracle.Params(beta, nk, 0)
        EC, ED, ED2 = P.table("EC"), P.table("ED"), P.table("ED2")
        n = np.arange(2, nk)
        # creation at an end strand of n (Z >= 2) vs destruction of the (1, n-1) pair (Z+1 >= 3)
        np.testing.assert_allclose(EC[n] / ED[n - 1], 1.0 / beta, rtol=3e-7)
        # Z = 1 -> 2: 1/(beta Nk) vs 1/(0.75 + (Nk-1) + 0.25)
        assert (1.0 / (beta * nk)) / ED2[nk - 1] == pytest.approx(1.0 / beta, rel=1e-6)


# ------------------------------------------------------------------ equilibrium init (App. C1-C3)
@pytest.mark.parametrize("beta", [1.0, 0.5, 2.0])
def test_init_Z_binomial(oracle, beta):
    nk, n = 25, 20000
    P = oracle.Params(beta, nk, 0)
    st = oracle.Ensemble(P, seed=9, n=n).state()
    Z = st["Z"]
    p = 1.0 / (1.0 + beta)
    k = np.arange(0, nk)
    pmf = stats.binom.pmf(k, nk - 1, p)
    obs = np.bincount(Z - 1, minlength=nk)[:nk]
    sel = pmf * n > 5
    chi = ((obs[sel] - n * pmf[sel]) ** 2 / (n * pmf[sel])).sum()
    assert stats.chi2.sf(chi, sel.sum() - 1) > 1e-4
    assert Z.mean() == pytest.approx(1 + (nk - 1) * p, abs=4 * math.sqrt((nk - 1) * p * (1 - p) / n))


def test_init_N_given_Z_uniform_composition(oracle):
    nk, n = 25, 20000
    P = oracle.Params(1.0, nk, 0)
    st = oracle.Ensemble(P, seed=4, n=n).state()
    Z, N = st["Z"], st["N"]
    for z in (8, 13, 18):
        rows = N[Z == z][:, :z]
        assert np.all(rows.sum(axis=1) == nk) and np.all(rows >= 1)
        # marginal of one part: C(Nk-n-1, Z-2)/C(Nk-1, Z-1) (common/chain.cu:60-88)
        vals = rows[:, 1:].ravel()
        m = np.arange(1, nk - z + 2)
        pmf = special.comb(nk - m - 1, z - 2) / special.comb(nk - 1, z - 1)
        obs = np.bincount(vals, minlength=len(m) + 1)[1:len(m) + 1]
        exp = pmf * len(vals)
        sel = exp > 5
        chi = ((obs[sel] - exp[sel]) ** 2 / exp[sel]).sum()
        assert stats.chi2.sf(chi, sel.sum() - 1) > 1e-5, z


def test_init_Q_gaussian_and_ends(oracle):
    nk, n = 60, 5000
    P = oracle.Params(1.0, nk, 1)
    st = oracle.Ensemble(P, seed=8, n=n).state()
    Z, N, Q = st["Z"], st["N"], st["Q"]
    idx = np.arange(N.shape[1])[None, :]
    internal = (idx > 0) & (idx < Z[:, None] - 1)
    scaled = (Q / np.sqrt(np.maximum(N, 1) / 3.0)[..., None])[internal].ravel()
    assert abs(scaled.mean()) < 0.02 and scaled.var() == pytest.approx(1.0, abs=0.02)
    assert stats.kstest(scaled, "norm").pvalue > 1e-4
    assert np.all(Q[:, 0] == 0)
    assert np.all(Q[np.arange(n), Z - 1] == 0)
    # initial link rates come from f_t: mean of the rate is <1/tau>
    w = st["w"][idx < Z[:, None] - 1]
    assert w.mean() == pytest.approx(P.scalars()["normdt"], rel=0.03)


# ------------------------------------------------------------------ dynamics
@pytest.mark.parametrize("cd,beta,nk", [(0, 1.0, 25), (1, 1.0, 40), (1, 2.0, 30)])
def test_events_keep_invariants_and_cache(oracle, cd, beta, nk):
    P = oracle.Params(beta, nk, cd)
    E = oracle.Ensemble(P, seed=3, n=300)
    for _ in range(5):
        E.advance(max_events=400)
        assert E.invariant_violations() == 0
        assert E.cache_error() == 0.0
    E64 = oracle.Ensemble(P, seed=3, n=300, fp64=True)
    E64.advance(max_events=800)
    assert E64.invariant_violations() == 0


def test_flow_events_keep_invariants(oracle):
    P = oracle.Params(1.0, 30, 1, kappa=[0, 0.1, 0, 0, 0, 0, 0, 0, 0])
    E = oracle.Ensemble(P, seed=3, n=200)
    E.advance(t_end=50.0)
    assert E.invariant_violations() == 0
    assert E.cache_error() == 0.0


def _stationary_check(oracle, beta, nk, cd, n, t):
    P = oracle.Params(beta, nk, cd)
    E = oracle.Ensemble(P, seed=21, n=n)
    E.advance(t_end=t)
    st = E.state()
    Z, N, Q = st["Z"], st["N"], st["Q"]
    p = 1.0 / (1.0 + beta)
    pmf = stats.binom.pmf(np.arange(nk), nk - 1, p)
    obs = np.bincount(Z - 1, minlength=nk)[:nk]
    sel = pmf * n > 5
    chi = ((obs[sel] - n * pmf[sel]) ** 2 / (n * pmf[sel])).sum()
    idx = np.arange(N.shape[1])[None, :]
    internal = (idx > 0) & (idx < Z[:, None] - 1)
    q2n = ((Q ** 2).sum(axis=2) / np.maximum(N, 1))[internal]
    return chi, sel.sum() - 1, q2n.mean(), q2n.size


@pytest.mark.parametrize("beta,nk,cd", [(1.0, 25, 0), (1.0, 25, 1), (0.5, 20, 1)])
def test_stationarity_of_equilibrium(oracle, beta, nk, cd):
    """A long run from the equilibrium draw keeps P(Z) binomial and <Q^2/N> = 1."""
    chi, dof, q2n, m = _stationary_check(oracle, beta, nk, cd, 6000, 300.0)
    assert stats.chi2.sf(chi, dof) > 1e-4
    assert q2n == pytest.approx(1.0, abs=5 * math.sqrt(2.0 / 3.0 / m) + 0.01)


def test_g0_equals_mean_internal_strands(oracle):
    """G(0) = <sigma_xy^2> = E[max(Z-2, 0)] at equilibrium (each internal strand gives 1)."""
    nk, n = 25, 4000
    P = oracle.Params(1.0, nk, 0)
    E = oracle.Ensemble(P, seed=13, n=n)
    _, s = E.advance(t_end=40.0, delta=1.0, k0=0, k_end=40, samples=True)
    g0 = (s.astype(np.float64) ** 2).mean()
    p = 0.5
    ez = 1 + (nk - 1) * p
    pz1 = (1 - p) ** (nk - 1)
    assert g0 == pytest.approx(ez - 2 + pz1, rel=0.03)


# ------------------------------------------------------------------ correlator (App. C7)
def test_correlator_level0_and_blocks_match_brute_force(oracle):
    p, m, s, n = 8, 2, 4, 3
    T = 200
    rng = np.random.default_rng(0)
    x = np.zeros((T, n, 3))
    for t in range(1, T):
        x[t] = 0.8 * x[t - 1] + rng.standard_normal((n, 3))
    cr = oracle.Correlator(p, m, s, n)
    for t in range(T):
        cr.add_row(x[t])
    lags, g = cr.result()
    # level 0: sum_t x_t . x_{t-j} / count / 3 / nchains
    for j in range(p):
        prod = (x[j:] * x[:T - j]).sum(axis=2).sum(axis=1)
        assert g[j] == pytest.approx(prod.sum() / (T - j) / 3 / n, rel=1e-12)
    # level 1: the same on block averages of m samples, lags j*m for j >= p/m
    xb = x[:T - T % m].reshape(-1, m, n, 3).mean(axis=1)
    Tb = xb.shape[0]
    for q, j in enumerate(range(p // m, p)):
        prod = (xb[j:] * xb[:Tb - j]).sum(axis=2).sum(axis=1)
        assert lags[p + q] == j * m
        assert g[p + q] == pytest.approx(prod.sum() / (Tb - j) / 3 / n, rel=1e-12)


def test_correlator_level_count(oracle):
    # common/main_cuda.cu:70 -- s = ceil(log(T/dt/p)/log(m)) + 1
    assert oracle.corr_levels(1e5, 1.0, 16, 2) == 14
    assert oracle.corr_levels(1e7, 1.0, 16, 2) == 21
    assert oracle.corr_levels(1000, 1.0, 64, 8) == 3


def test_fp64_mode_agrees_statistically(oracle):
    """FP32 mirror and FP64 oracles share streams; their ensemble means agree."""
    nk, n = 25, 3000
    P = oracle.Params(1.0, nk, 1)
    a = oracle.Ensemble(P, seed=2, n=n)
    b = oracle.Ensemble(P, seed=2, n=n, fp64=True)
    ea, sa = a.advance(t_end=60.0, delta=1.0, k0=0, k_end=60, samples=True)
    eb, sb = b.advance(t_end=60.0, delta=1.0, k0=0, k_end=60, samples=True)
    assert ea == pytest.approx(eb, rel=0.01)
    za, zb = a.state()["Z"].mean(), b.state()["Z"].mean()
    assert za == pytest.approx(zb, abs=0.25)
    assert (sa.astype(np.float64) ** 2).mean() == pytest.approx((sb ** 2).mean(), rel=0.05)


def test_flow_samples_stay_inside_the_window(oracle):
    """A flow-output ensemble without shear starts its sample counter before the window;
    samples before k0 are consumed without being written (guard rows stay untouched)."""
    import ctypes as C
    P = oracle.Params(1.0, 24, 1, kappa=[0] * 9)
    E = oracle.Ensemble(P, seed=40, n=37, salt=0)
    k0, k_end, n = 1, 5, 37
    buf = np.full((k_end - k0 + 2, n, 8), np.float32(-7.0), np.float32)   # one guard row each side
    oracle.lib().orc_params_set_flow_samples(P.h, 1)
    inner = buf[1:-1]
    assert inner.flags["C_CONTIGUOUS"]
    oracle.lib().orc_ens_advance(E.h, 12.0, 2 ** 62, 3.0, k0, k_end, inner.ctypes.data_as(C.c_void_p), 2)
    assert np.all(buf[0] == -7.0) and np.all(buf[-1] == -7.0)
    assert np.all(inner[:, :, 7] >= 1)          # every row written (Z column)


def test_long_horizon_gt_runner_matches_ensemble_path(oracle):
    """orc_gt_chains (one chain at a time, windows of samples fed to a per-chain C
    correlator, nothing stored) gives the same G(t) as the ensemble path's samples fed
    to the brute-force-pinned numpy correlator, for any window length."""
    P = oracle.Params(1.0, 25, 0)
    n, T = 64, 300.0
    s = oracle.corr_levels(T, 1.0, 16, 2)
    E = oracle.Ensemble(P, seed=1, n=n)
    _, smp = E.advance(t_end=T, delta=1.0, k0=0, k_end=int(T), samples=True)
    cr = oracle.Correlator(16, 2, s, n)
    for row in smp:
        cr.add_row(row)
    lag_o, g_o = cr.result()
    for window, fp64 in ((37, False), (256, False), (64, True)):
        lags, G, ev, _ = oracle.gt_chains(P, 1, n, T, s=s, fp64=fp64, window=window)
        assert np.array_equal(lags, lag_o)
        if not fp64:
            assert ev == int(E.state()["ev"].sum())
            np.testing.assert_allclose(G.mean(axis=0), g_o, rtol=1e-12, atol=1e-12)
        else:
            assert abs(G.mean(axis=0)[0] - g_o[0]) < 0.01 * g_o[0]
