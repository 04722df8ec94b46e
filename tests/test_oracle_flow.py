"""Pins for the oracle's flow path (CPU only): affine deformation, full stress tensor and
Lpp against the mathematics, independently of the formulas the oracle itself uses.

Reference semantics (common/ensemble_kernel.cu):
  * deformation, explicit Euler, kappa = grad v^T row-major (:105-110): Q_a += dt * k_ab Q_b;
  * the step's rates use Q deformed with the PREVIOUS step's dt (:135-146, tdt[j]); the
    stress output applies the pending deformation with that dt as well (:512-522);
  * stress per chain sigma_ab = -sum_i 3 Q_ia Q_ib / N_i, Lpp = sum_i |Q_i|, Z (:504-538).

Each test below is chosen so that a plausible mistake fails it:
  * exact affine map from the clock's own dt sequence -> transposed kappa, a missing or
    extra deformation step, deformation with another step's dt, kappa not rounded as the
    reference's float constants;
  * closed forms at rest -> stress sign, the factor 3, the division by N, Lpp (|Q|, not
    |Q|^2), the Z column;
  * linear response (start-up shear and uniaxial extension against -rate * int G from
    the same chains' equilibrium stress autocorrelation) -> sign and magnitude of the
    coupling between deformation and stress over time;
  * Green-Kubo steady shear (eta0 = int G) and the second-order normal stress
    difference N1 = -2 rate^2 int t G(t) dt -> the steady state, and (N1's sign) a
    transposed kappa in the dynamics.
"""
import math

import numpy as np
import pytest


def _internal_mean(nk, beta, f):
    """E[sum over internal strands of f(N)] at equilibrium: each of the nk-1 inner
    Kuhn-step boundaries is an entanglement with probability p = 1/(1+beta); an
    internal strand of n Kuhn steps spans two entangled boundaries with none between
    (nk-1-n placements)."""
    p = 1.0 / (1.0 + beta)
    return sum((nk - 1 - n) * p * p * (1.0 - p) ** (n - 1) * f(n) for n in range(1, nk - 1))


def _flow_samples(oracle, nk, n, T, dlt, kappa, cd=0, beta=1.0, seed=9, fp64=True):
    P = oracle.Params(beta, nk, cd, kappa=kappa)
    E = oracle.Ensemble(P, seed=seed, n=n, fp64=fp64)
    K = int(round(T / dlt)) + 1
    _, s = E.advance(t_end=T, delta=dlt, k0=0, k_end=K, samples=True, flow_samples=True)
    return s.astype(np.float64)  # [sample k at time k*dlt][chain][xx yy zz xy yz zx Lpp Z]


KAPPA = np.array([0.013, 0.031, -0.007, 0.002, -0.011, 0.019, -0.023, 0.005, -0.002])


@pytest.mark.parametrize("fp64", [True, False], ids=["fp64", "fp32"])
def test_affine_map_follows_the_clock(oracle, fp64):
    """Shifts move Kuhn steps between strands and leave every strand vector in place, so
    a chain that only shifted (Z unchanged after every single jump; CD off) must hold
    Q_k = prod_j (I + dt_j K) Q_0 over the dt of every jump so far, dt_j read from the
    chain's own clock.  K is a general non-symmetric matrix."""
    nk, n, jumps = 25, 3000, 12
    P = oracle.Params(1.0, nk, 0, kappa=KAPPA)
    E = oracle.Ensemble(P, seed=5, n=n, fp64=fp64)
    s = E.state()
    qkey = "Q64" if fp64 else "Q"
    Z0, Q0, t_prev = s["Z"].copy(), s[qkey].astype(np.float64), s["t"].copy()
    K = KAPPA.astype(np.float32).astype(np.float64).reshape(3, 3)  # the reference's float constants
    eye = np.broadcast_to(np.eye(3), (n, 3, 3))
    M, MT, Mlag = eye.copy(), eye.copy(), eye.copy()  # reference; transposed K; last dt missing
    shift_only = np.ones(n, bool)
    for _ in range(jumps):
        E.advance(max_events=1)
        s = E.state()
        dt = s["t"] - t_prev
        t_prev = s["t"].copy()
        shift_only &= s["Z"] == Z0
        Mlag = M.copy()
        M = M + dt[:, None, None] * np.einsum("ij,njk->nik", K, M)
        MT = MT + dt[:, None, None] * np.einsum("ij,njk->nik", K.T, MT)
    sel = np.nonzero(shift_only)[0]
    assert sel.size >= 50, "too few shift-only chains (%d)" % sel.size
    idx = np.arange(Q0.shape[1])[None, :]
    internal = ((idx > 0) & (idx < Z0[:, None] - 1))[sel]

    def mapped(m):
        return np.einsum("nij,nsj->nsi", m[sel], Q0[sel])[internal]

    want = mapped(M)
    got = s[qkey].astype(np.float64)[sel][internal]
    tol = (1e-12 if fp64 else 2e-5) * np.max(np.abs(want))
    assert np.max(np.abs(got - want)) <= tol
    # the mistakes this pins are far outside the tolerance on the same chains
    assert np.max(np.abs(mapped(MT) - want)) > 100 * tol
    assert np.max(np.abs(mapped(Mlag) - want)) > 100 * tol


@pytest.mark.parametrize("fp64", [True, False], ids=["fp64", "fp32"])
def test_stress_at_rest_closed_forms(oracle, fp64):
    """At equilibrium (kappa = 0) every internal strand has Q ~ N(0, N/3 I), so
    <sigma_aa> = -E[#internal] = -E[max(Z-2, 0)] for a = x, y, z, <sigma_ab> = 0 for
    a != b, <Lpp> = E[sum_internal 2 sqrt(2N/(3 pi))] and <Z> = 1 + (Nk-1)/(1+beta);
    checked at t = 0 and after 20 tau_K of dynamics (stationarity)."""
    nk, beta, n = 25, 1.0, 20000
    s = _flow_samples(oracle, nk, n, 20.0, 20.0, [0] * 9, beta=beta, fp64=fp64)
    p = 1.0 / (1.0 + beta)
    g0 = _internal_mean(nk, beta, lambda m: 1.0)
    assert g0 == pytest.approx(1 + (nk - 1) * p - 2 + (1 - p) ** (nk - 1), rel=1e-12)
    lpp = _internal_mean(nk, beta, lambda m: 2.0 * math.sqrt(2.0 * m / (3.0 * math.pi)))
    want = [-g0, -g0, -g0, 0.0, 0.0, 0.0, lpp, 1 + (nk - 1) * p]
    for k in range(s.shape[0]):
        mean = s[k].mean(axis=0)
        se = s[k].std(axis=0) / math.sqrt(n)
        for q in range(8):
            assert abs(mean[q] - want[q]) < 5 * se[q] + 1e-9, (k, q, mean[q], want[q], se[q])


def _linear_response(oracle, nk, n, T, dlt, rate, uniaxial):
    """Per-chain X_c(t) = A_c(t) + fac*rate*int_0^t B_c: A = response of the sheared (or
    stretched) run minus the run at rest on the same chains and Philox streams; B = the
    rest run's shear-stress autocorrelation (xy, yz, xz, time origin 0).  Linear
    response: E[A(t)] = -fac * rate * int_0^t G(s) ds with G = E[B] (fac 1 for sigma_xy
    in shear, 3 for sigma_xx - sigma_yy in uniaxial extension)."""
    kap = [rate, 0, 0, 0, -rate / 2, 0, 0, 0, -rate / 2] if uniaxial else [0, rate, 0, 0, 0, 0, 0, 0, 0]
    s0 = _flow_samples(oracle, nk, n, T, dlt, [0] * 9)
    s1 = _flow_samples(oracle, nk, n, T, dlt, kap)
    off = s0[:, :, 3:6]
    B = (off * off[0:1]).mean(axis=2)
    if uniaxial:
        A = (s1[:, :, 0] - s1[:, :, 1]) - (s0[:, :, 0] - s0[:, :, 1])
        fac = 3.0
    else:
        A = s1[:, :, 3] - s0[:, :, 3]
        fac = 1.0
    intB = np.concatenate([np.zeros((1, n)), np.cumsum(0.5 * (B[1:] + B[:-1]) * dlt, axis=0)])
    return A, fac * rate * intB, s0, s1


@pytest.mark.parametrize("uniaxial", [False, True], ids=["shear", "uniaxial"])
def test_startup_follows_linear_response(oracle, uniaxial):
    """Start-up from the equilibrium draw at a small rate (1e-3 per tau_K, Nk = 25):
    sigma_xy(t) = -rate int_0^t G (shear), sigma_xx - sigma_yy = -3 rate int_0^t G
    (uniaxial, Trouton), for t in [2, 10] tau_K; below ~2 tau_K the sample's own clock
    offset (a sample is taken at the first jump past k*delta, the deformation runs to
    that jump) is not negligible against the integral."""
    nk, n, T, dlt, rate = 25, 20000, 10.0, 0.25, 1e-3
    A, R, _, _ = _linear_response(oracle, nk, n, T, dlt, rate, uniaxial)
    X = A + R
    for k in range(int(2.0 / dlt), A.shape[0]):
        pred = -R[k].mean()
        se = X[k].std() / math.sqrt(n)
        assert pred < 0
        assert abs(A[k].mean() - pred) < 4 * se + 0.02 * abs(pred), (k * dlt, A[k].mean(), pred, se)
    # at early times the response is the affine one: -fac rate G(0) t
    assert A[int(2.0 / dlt)].mean() < 0


def test_green_kubo_and_second_normal_stress_difference(oracle):
    """Steady shear at Weissenberg < 1 (Nk = 8, rate 1e-2, 100 tau_K, terminal time a few
    tau_K): the steady shear stress is -rate * eta0 with eta0 = int_0^inf G (Green-Kubo),
    and the first normal stress difference is second order, N1 = -2 rate^2 int t G(t) dt
    in the reference's sign convention (sigma = -sum 3QQ/N).  A transposed kappa leaves
    sigma_xy unchanged but flips the sign of N1."""
    nk, n, T, dlt, rate = 8, 60000, 100.0, 1.0, 1e-2
    A, R, s0, s1 = _linear_response(oracle, nk, n, T, dlt, rate, False)
    K = A.shape[0]
    last = slice(K // 2, K)
    a_c, r_c = A[last].mean(axis=0), R[last].mean(axis=0)
    se = (a_c + r_c).std() / math.sqrt(n)
    assert abs(a_c.mean() + r_c.mean()) < 4 * se + 0.02 * r_c.mean(), (a_c.mean(), -r_c.mean(), se)
    off = s0[:, :, 3:6]
    G = (off * off[0:1]).mean(axis=2).mean(axis=1)
    t = np.arange(K) * dlt
    psi = float(np.sum(0.5 * (t[1:] * G[1:] + t[:-1] * G[:-1]) * dlt))
    n1_c = ((s1[:, :, 0] - s1[:, :, 1]) - (s0[:, :, 0] - s0[:, :, 1]))[last].mean(axis=0)
    n1, n1_se = n1_c.mean(), n1_c.std() / math.sqrt(n)
    pred = -2.0 * rate * rate * psi
    assert n1 < -5 * n1_se, (n1, n1_se)
    assert abs(n1 - pred) < 4 * n1_se + 0.15 * abs(pred), (n1, pred, n1_se)
