"""Pins for the oracle's random numbers and FP32 math (CPU only).

Philox4x32-10 is pinned by the published Random123 known-answer vectors and by
vectors from rocRAND's independent implementation; exp/log/sin/cos by error
bounds against libm in double; the scan tree by exact prefix sums.
"""
import json
import math
import os

import numpy as np
import pytest
from scipy import stats

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


# Random123 kat_vectors for philox4x32-10 (Salmon et al., SC'11 reference code)
RANDOM123_KAT = [
    ([0, 0, 0, 0], [0, 0], [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
    ([0xffffffff] * 4, [0xffffffff] * 2, [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
    ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0],
     [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
]


@pytest.mark.parametrize("ctr,key,out", RANDOM123_KAT)
def test_philox_known_answers(oracle, ctr, key, out):
    assert oracle.philox(ctr, key) == out


def test_philox_matches_rocrand(oracle):
    vec = json.load(open(os.path.join(GOLDEN, "philox_rocrand.json")))
    assert len(vec) == 100
    for v in vec:
        assert oracle.philox(v["ctr"], v["key"]) == v["out"], v


def test_u01_open_interval(oracle):
    assert oracle.u01(0) == 2.0 ** -24
    assert oracle.u01(0xFFFFFFFF) == 1.0 - 2.0 ** -24
    assert oracle.u01(0x80000000) == 0.5 + 2.0 ** -24
    # 9 low bits are dropped; every value is exact
    assert oracle.u01(0x12345600) == oracle.u01(0x123457FF)
    ks = np.arange(0, 1 << 23, 4099, dtype=np.uint64)
    vals = np.array([oracle.u01(int(k) << 9) for k in ks])
    assert np.array_equal(vals, (ks + 0.5) * 2.0 ** -23)


def _ulp_err(got, ref):
    got = np.asarray(got, np.float64)
    ref = np.asarray(ref, np.float64)
    ulp = np.spacing(np.abs(ref).astype(np.float32)).astype(np.float64)
    return np.abs(got - ref) / ulp


def test_exp_error_bound(oracle):
    x = np.linspace(-86.9, 87.9, 400001).astype(np.float32)
    got = oracle.vec_math("exp", x)
    ref = np.exp(x.astype(np.float64))
    assert _ulp_err(got, ref).max() <= 2.0
    assert oracle.expf(0.0) == 1.0
    assert oracle.expf(-90.0) == 0.0 and math.isinf(oracle.expf(89.0))


def test_log_error_bound(oracle):
    u = ((np.arange(0, 1 << 24, 31) + 0.5) * 2.0 ** -24).astype(np.float32)
    wide = np.exp(np.linspace(-85, 85, 200001)).astype(np.float32)
    for x in (u, wide):
        got = oracle.vec_math("log", x)
        ref = np.log(x.astype(np.float64))
        err = np.abs(got - ref)
        # relative bound away from 1, absolute bound near 1 (ln x -> 0)
        assert np.all(err <= 2.5 * np.spacing(np.maximum(np.abs(ref), 1e-7).astype(np.float32)) + 1.2e-7 * np.abs(x - 1.0) + 1e-12)
    assert oracle.logf(1.0) == 0.0


def test_sincos_error_bound(oracle):
    v = ((np.arange(0, 1 << 24, 17) + 0.5) * 2.0 ** -24).astype(np.float32)
    s = oracle.vec_math("sin2pi", v)
    c = oracle.vec_math("cos2pi", v)
    th = 2 * np.pi * v.astype(np.float64)
    assert np.abs(s - np.sin(th)).max() < 3e-7
    assert np.abs(c - np.cos(th)).max() < 3e-7
    assert np.abs(s.astype(np.float64) ** 2 + c.astype(np.float64) ** 2 - 1).max() < 5e-7


def test_pow_error(oracle):
    rng = np.random.default_rng(1)
    for _ in range(2000):
        x = float(np.float32(rng.uniform(0.3, 20.0)))
        y = float(np.float32(rng.uniform(-4.0, -0.5)))
        ref = x ** y
        assert abs(oracle.powf(x, y) - ref) <= 2e-6 * ref


def test_boxmuller_is_standard_normal(oracle):
    n = 40000
    g = []
    for i in range(n // 2):
        x = oracle.philox([i, 0, 77, oracle.PURPOSE_GAUSS], [1, 2])
        g.extend(oracle.boxmuller(x[0], x[1]))
    g = np.array(g)
    assert abs(g.mean()) < 4 / math.sqrt(n)
    assert abs(g.var() - 1) < 0.03
    assert stats.kstest(g, "norm").pvalue > 1e-3


def test_scan_tree_is_prefix_sum(oracle):
    rng = np.random.default_rng(3)
    for n in (1, 5, 63, 64, 65, 128, 200, 640):
        v = rng.random(n).astype(np.float32)
        incl = oracle.scan_tree(v)
        exact = np.cumsum(v.astype(np.float64))
        assert np.all(np.abs(incl - exact) <= 8 * np.spacing(exact.astype(np.float32)))
        # zero rates add exactly nothing
        assert incl[0] == v[0]
    z = np.zeros(100, np.float32)
    z[37] = 1.5
    inc = oracle.scan_tree(z)
    assert np.all(inc[:37] == 0) and np.all(inc[37:] == 1.5)


def test_scan_tree_exact_on_integers_and_total(oracle):
    """Small integers sum exactly in any order: every inclusive sum and the total are
    the exact prefix sums, over block boundaries (64) and ragged tails."""
    rng = np.random.default_rng(4)
    for n in (1, 2, 4, 5, 63, 64, 65, 127, 128, 129, 256, 300, 1024):
        v = rng.integers(0, 1000, size=n).astype(np.float32)
        incl = oracle.scan_tree(v)
        assert np.array_equal(incl, np.cumsum(v.astype(np.int64)).astype(np.float32))
        assert oracle.tree_total(v) == float(v.astype(np.int64).sum())


def test_scan_tree_order_is_blockwise(oracle):
    """The association is lane-serial / 16-lane Hillis-Steele / block-serial: a value
    pattern whose FP32 sum depends on the order separates it from a plain running sum."""
    big, small = np.float32(2 ** 24), np.float32(1.0)
    # running sum: big + 1 + 1 + 1 = big (each +1 rounds away); lane-serial over
    # slots 0..3 of one lane gives the same; but Hillis-Steele adds lane 1's own
    # (1+1+1+1 = 4) to lane 0's total, which is exact.
    v = np.array([big, 0, 0, 0, small, small, small, small], np.float32)
    incl = oracle.scan_tree(v)
    run = np.float32(0)
    seq = []
    for x in v:
        run = np.float32(run + x)
        seq.append(run)
    assert seq[-1] == big            # sequential sum loses the small terms
    assert incl[-1] == big + 4       # the tree adds them up first
    assert incl[4] == big            # inside lane 1: X + s_0 = big + 1 rounds to big
