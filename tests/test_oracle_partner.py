"""Pins for the polydisperse CD background (CPU only).

With PD_flag = 1 the reference draws the length of the chain on the other side of each
new entanglement from a GEX weight distribution (common/gamma.cpp:174-263,
common/gpu_random.cu:62-76,212-228, common/chain.cu:141-149) and averages the CD creation
prefactor over it (common/ensemble.cu:137-146).  The oracle holds the distribution as
65536 equiprobable quantiles; a second table re-weights them by the mean destruction rate
for CD-created links.  The polydisperse-melt mode uses the ensemble's own chain lengths.
"""
import math

import numpy as np
import pytest
from scipy import integrate, special

from test_oracle_model import _pcd_double

K = 65536
GEX = (1.5, 1.0, 40000.0, 120.0)   # a, b, mp, Mk: partner Nk = m * 333.3


def _normdt(beta, nk):
    """<1/tau> of p_CD(beta, Nk) in double (common/pcd_tau.h:36-66)."""
    c = _pcd_double(beta, nk)
    a, t0, tm, td, g = c["alpha"], c["tau_0"], c["tau_max"], c["tau_d"], c["g"]
    c1 = tm ** a - t0 ** a
    cm1 = tm ** (a - 1) - t0 ** (a - 1)
    rc = (a - 1) / a / t0 if c1 == 0 else cm1 / c1
    return (1 - g) * a / (a - 1) * rc + g / td


@pytest.mark.parametrize("s", [0.7, 2.5, 11.0])
def test_incomplete_gamma(oracle, s):
    for x in (1e-3, 0.3, s * 0.9, s + 1.0, 2 * s + 3, 40.0):
        assert oracle.lib().orc_gamma_p(s, x) == pytest.approx(special.gammainc(s, x), rel=1e-12, abs=1e-15)


def test_gex_quantiles_invert_the_weight_distribution(oracle):
    a, b, mp, Mk = GEX
    P = oracle.Params(1.0, 100, 1, partner="gex", gex=GEX)
    nk = P.table("PF_NK").astype(np.float64)
    assert nk.size == K and np.all(np.diff(nk) >= 0)
    m = nk * Mk / mp
    prob = special.gammainc((a + 1) / b, m ** b)
    np.testing.assert_allclose(prob, (np.arange(K) + 0.5) / K, atol=2e-6)


def test_cd_created_partners_are_weighted_by_destruction_rate(oracle):
    """The f_d_t table holds the same lengths re-weighted by <1/tau>(Nk): its mean of any
    function of Nk equals the normdt-weighted mean over the f_t table (to 1/K)."""
    beta = 1.0
    P = oracle.Params(beta, 100, 1, partner="gex", gex=GEX)
    f = P.table("PF_NK").astype(np.float64)
    d = P.table("PD_NK").astype(np.float64)
    w = np.array([_normdt(beta, x) for x in f])
    for g in (np.log, np.sqrt, lambda x: x):
        assert np.mean(g(d)) == pytest.approx(np.sum(w * g(f)) / np.sum(w), rel=2e-3)
    # and differs from the unweighted mean (short partners release entanglements faster)
    assert np.mean(d) < 0.99 * np.mean(f)


def test_creation_prefactor_is_the_background_average(oracle):
    """c = <normdt>/beta averaged over the GEX weight distribution, by quadrature."""
    a, b, mp, Mk = GEX
    beta = 1.0
    P = oracle.Params(beta, 100, 1, partner="gex", gex=GEX)
    s = (a + 1) / b
    dens = lambda m: b * m ** a * math.exp(-m ** b) / math.gamma(s)
    ref, _ = integrate.quad(lambda m: dens(m) * _normdt(beta, m * mp / Mk), 0, 60, limit=400)
    assert P.scalars()["c_create"] == pytest.approx(ref / beta, rel=2e-3)
    # CC table follows the new prefactor
    cc = P.table("CC")
    assert cc[5] == pytest.approx(4 * P.scalars()["c_create"], rel=1e-6)


def test_melt_partners_follow_the_length_weight_distribution(oracle):
    rng = np.random.default_rng(4)
    nk_chains = rng.choice([20, 50, 80, 200], size=1000, p=[0.4, 0.3, 0.2, 0.1]).astype(np.int32)
    P = oracle.Params(1.0, 200, 1, partner="melt", nk_chains=nk_chains)
    tab = P.table("PF_NK")
    vals, cnt = np.unique(nk_chains, return_counts=True)
    wfrac = vals * cnt / np.sum(vals * cnt)
    got = np.array([(tab == v).mean() for v in vals])
    np.testing.assert_allclose(got, wfrac, atol=1.5 / K)


def test_initial_links_draw_from_the_mixture(oracle):
    """Initial CD rates come from f_t of a partner drawn per link: their mean is the
    background average of <1/tau>, not the chain's own."""
    beta, nk, n = 1.0, 60, 4000
    P = oracle.Params(beta, nk, 1, partner="gex", gex=GEX)
    E = oracle.Ensemble(P, seed=9, n=n)
    st = E.state()
    idx = np.arange(st["N"].shape[1])[None, :]
    w = st["w"][idx < st["Z"][:, None] - 1].astype(np.float64)
    mix = np.mean([_normdt(beta, x) for x in P.table("PF_NK").astype(np.float64)])
    assert w.mean() == pytest.approx(mix, rel=0.03)
    assert abs(w.mean() - _normdt(beta, nk)) > 0.1 * mix
