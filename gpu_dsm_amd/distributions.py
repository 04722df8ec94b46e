"""Z / N / Q distributions of an ensemble's state and their equilibrium laws.

The reference writes three distributions with `-distr` (common/ensemble.cu:356-484):
the pdf of Z over chains, the pdf of N over all strands, and the CDF of |Q| over the
internal strands.  At equilibrium the slip-link model fixes all three (independent
entanglements at the Nk-1 inner Kuhn-step boundaries with probability p = 1/(1+beta),
Gaussian strand vectors of variance N/3 per component):

  * Z - 1 ~ Binomial(Nk - 1, p) per chain;
  * the expected number of internal strands of length n in a chain is
    (Nk - 1 - n) p^2 (1 - p)^(n-1)  (two entanglements n boundaries apart, none between);
  * 3 |Q|^2 / N of an internal strand ~ chi-squared with 3 degrees of freedom.

summarize() reports the observed distributions against these laws in units of their
sampling error, so a run that drifts from equilibrium (or a state with a broken move)
shows up as many-sigma deviations.
"""
import ctypes as C
import math

import numpy as np

from . import engine as eng


def fetch_zn_q(e, ens):
    """Z[n], N[n, cap] and Q[n, cap, 3] of an ensemble (no other state arrays)."""
    info = e.ens[ens]
    n, cap = info["n"], info["cap"]
    Z = np.zeros(n, np.int32)
    N = np.zeros((n, cap), np.int32)
    Q = np.zeros((n, cap, 3), np.float32)
    eng._check(e._L.dsm_get_state(e.h, int(ens), int(cap), eng._p(Z, C.c_int32), eng._p(N, C.c_int32),
                                  eng._p(Q, C.c_float), None, None, None, None, None, None))
    return Z, N, Q


def _erf(r):
    try:
        from scipy.special import erf
        return erf(r)
    except ImportError:  # the same function, element by element
        return np.vectorize(math.erf)(r)


def chi2_3_cdf(x):
    """CDF of the chi-squared law with 3 degrees of freedom: erf(sqrt(x/2)) -
    sqrt(2x/pi) exp(-x/2)."""
    x = np.maximum(np.asarray(x, np.float64), 0.0)
    return _erf(np.sqrt(x / 2.0)) - np.sqrt(2.0 * x / math.pi) * np.exp(-x / 2.0)


def internal_masks(Z, cap):
    """Boolean [n, cap]: strand slots 1..Z-2 (internal strands) and 0..Z-1 (all strands)."""
    k = np.arange(cap)[None, :]
    return (k >= 1) & (k < Z[:, None] - 1), k < Z[:, None]


def reference_files(Z, N, Q, max_q_rows=None):
    """The three -distr tables as the reference writes them: (Z, pdf), (N, pdf over all
    strands), (|Q| sorted over internal strands, CDF).  max_q_rows: keep that many evenly
    spaced rows of the |Q| table (the reference writes one row per internal strand)."""
    internal, alls = internal_masks(Z, N.shape[1])
    zv, zc = np.unique(Z, return_counts=True)
    nv, nc = np.unique(N[alls], return_counts=True)
    q = np.sort(np.sqrt(np.sum(Q[internal].astype(np.float64) ** 2, axis=1)))
    cdf = (np.arange(q.size) + 1.0) / max(q.size, 1)
    if max_q_rows and q.size > max_q_rows:
        keep = np.unique(np.linspace(0, q.size - 1, max_q_rows).astype(np.int64))
        q, cdf = q[keep], cdf[keep]
    return (zv, zc / Z.size), (nv, nc / max(nc.sum(), 1)), (q, cdf)


def summarize(Z, N, Q, nk, beta, n_bins=12):
    """Observed distributions against the equilibrium laws.  nk[n]: each chain's Nk."""
    nk = np.asarray(nk, np.int64)
    p = 1.0 / (1.0 + beta)
    n = Z.size
    # Z: standardized per-chain residual (mean 0, variance 1 at equilibrium)
    mu = 1.0 + (nk - 1) * p
    sd = np.sqrt(np.maximum((nk - 1) * p * (1.0 - p), 1e-300))
    u = (Z - mu) / sd
    z_out = dict(mean_u=float(u.mean()), mean_u_sigma=float(u.mean() * math.sqrt(n)),
                 var_u=float(u.var()), var_u_sigma=float((u.var() - 1.0) / math.sqrt(2.0 / n)),
                 mean_Z=float(Z.mean()), expected_mean_Z=float(mu.mean()))
    # internal N: counts per length against the expected counts
    internal, _ = internal_masks(Z, N.shape[1])
    Ni = N[internal]
    lengths = np.arange(1, n_bins + 1)
    obs = np.array([(Ni == m).sum() for m in lengths], np.float64)
    exp = np.array([np.sum(np.maximum(nk - 1 - m, 0) * p * p * (1.0 - p) ** (m - 1)) for m in lengths])
    dev = (obs - exp) / np.sqrt(np.maximum(exp, 1e-300))
    n_out = dict(lengths=lengths.tolist(), observed=obs.astype(np.int64).tolist(), expected=exp.tolist(),
                 max_abs_dev_sigma=float(np.max(np.abs(dev))), chi2_per_bin=float(np.mean(dev ** 2)),
                 internal_strands=int(Ni.size))
    # Q: 3|Q|^2/N of internal strands against chi-squared(3)
    x = 3.0 * np.sum(Q[internal].astype(np.float64) ** 2, axis=1) / Ni
    x.sort()
    F = chi2_3_cdf(x)
    m = x.size
    ks = float(max(np.max(np.arange(1, m + 1) / m - F), np.max(F - np.arange(m) / m))) if m else 0.0
    q_out = dict(ks_chi2_3=ks, ks_sqrt_n=ks * math.sqrt(m), mean_3q2_over_n=float(x.mean()) if m else 0.0)
    return dict(chains=int(n), Z=z_out, N=n_out, Q=q_out)


def two_sample(a_state, b_state, n_bins=12):
    """Two states of the same chains (for instance from FP32 and FP64 arithmetic): the
    largest difference of the Z and internal-N pmfs and of the 3|Q|^2/N CDF."""
    (Za, Na, Qa), (Zb, Nb, Qb) = a_state, b_state
    zmax = int(max(Za.max(), Zb.max()))
    pz = lambda Z: np.bincount(Z, minlength=zmax + 1) / Z.size
    dz = float(np.max(np.abs(pz(Za) - pz(Zb))))
    ia, _ = internal_masks(Za, Na.shape[1])
    ib, _ = internal_masks(Zb, Nb.shape[1])
    pn = lambda v: np.bincount(np.minimum(v, n_bins + 1), minlength=n_bins + 2)[1:] / max(v.size, 1)
    dn = float(np.max(np.abs(pn(Na[ia]) - pn(Nb[ib]))))
    xa = np.sort(3.0 * np.sum(Qa[ia].astype(np.float64) ** 2, axis=1) / Na[ia])
    xb = np.sort(3.0 * np.sum(Qb[ib].astype(np.float64) ** 2, axis=1) / Nb[ib])
    grid = np.concatenate([xa, xb])
    ks = float(np.max(np.abs(np.searchsorted(xa, grid, side="right") / xa.size -
                             np.searchsorted(xb, grid, side="right") / xb.size)))
    return dict(max_dpmf_Z=dz, max_dpmf_N=dn, ks_3q2_over_n=ks,
                ks_scale=float(math.sqrt((xa.size + xb.size) / (xa.size * xb.size))))
