"""The equilibrium-law summaries of gpu_dsm_amd.distributions, on states the oracle
draws (CPU only).  At equilibrium every deviation is a few sampling errors at most;
planted changes of the kind a broken move would make (a Z offset, strands one Kuhn
step longer, strand vectors 5% too long) are many errors out."""
import numpy as np
import pytest

from gpu_dsm_amd import distributions as D


def _draw(oracle, n=3000, seed=4, beta=1.0):
    rng = np.random.default_rng(seed)
    nk = rng.choice([12, 40, 90, 200], size=n, p=[0.2, 0.3, 0.3, 0.2]).astype(np.int32)
    P = oracle.Params(beta, int(nk.max()), 1, nmax=int(nk.max()), partner="melt", nk_chains=nk)
    E = oracle.Ensemble(P, seed=seed, n=n, nk_per_chain=nk)
    s = E.state()
    return s["Z"], s["N"], s["Q"], nk


@pytest.mark.parametrize("beta", [0.5, 1.0, 2.0])
def test_equilibrium_draw_matches_the_laws(oracle, beta):
    Z, N, Q, nk = _draw(oracle, beta=beta)
    out = D.summarize(Z, N, Q, nk, beta)
    assert abs(out["Z"]["mean_u_sigma"]) < 4.5 and abs(out["Z"]["var_u_sigma"]) < 4.5, out["Z"]
    assert out["N"]["max_abs_dev_sigma"] < 4.5, out["N"]
    assert out["Q"]["ks_sqrt_n"] < 2.0, out["Q"]


def test_planted_deviations_are_caught(oracle):
    Z, N, Q, nk = _draw(oracle)
    bad_z = D.summarize(Z + 1, N, Q, nk, 1.0)
    assert abs(bad_z["Z"]["mean_u_sigma"]) > 10
    bad_n = D.summarize(Z, N + 1, Q, nk, 1.0)
    assert bad_n["N"]["max_abs_dev_sigma"] > 20
    bad_q = D.summarize(Z, N, Q * np.float32(1.05), nk, 1.0)
    assert bad_q["Q"]["ks_sqrt_n"] > 4


def test_two_sample_of_a_state_with_itself_and_a_planted_change(oracle):
    Z, N, Q, nk = _draw(oracle)
    same = D.two_sample((Z, N, Q), (Z, N, Q))
    assert same["max_dpmf_Z"] == 0 and same["max_dpmf_N"] == 0 and same["ks_3q2_over_n"] == 0
    moved = D.two_sample((Z, N, Q), (Z, N, Q * np.float32(1.05)))
    assert moved["ks_3q2_over_n"] > 4 * moved["ks_scale"]


def test_reference_tables(oracle):
    Z, N, Q, nk = _draw(oracle, n=500)
    (zv, zp), (nv, npdf), (q, cdf) = D.reference_files(Z, N, Q)
    assert abs(zp.sum() - 1) < 1e-12 and abs(npdf.sum() - 1) < 1e-12
    assert np.all(np.diff(q) >= 0) and cdf[-1] == 1.0
    assert q.size == int(np.sum(np.maximum(Z - 2, 0)))
