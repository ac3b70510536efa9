"""Pins the oracle's vector ops (nbody/nbody_engine.cpp:47-113,
nbody/nbody_engine_simple.cpp:237-419) with element-wise host expressions on
exactly representable inputs, the base-class corner cases and the eps/2
Kahan known-answer test (test/engine/test_nbody_engine.cpp:377-437)."""
import numpy as np
import pytest

import oracle_lib

P = oracle_lib.p
EPS = np.finfo(np.float64).eps


def _arr(vs):
    return (oracle_lib.PD * max(1, len(vs)))(*[P(v) for v in vs])


def _ints(rng, n, hi=10000):
    return rng.integers(0, hi, n).astype(np.float64)


def test_fmadd_exact_on_integers():
    rng = np.random.default_rng(1)
    n = 1000
    a, b, c = _ints(rng, n), _ints(rng, n), _ints(rng, n)
    x = a.copy()
    oracle_lib.lib().oracle_fmadd_inplace(P(x), P(b), 5.0, n)
    assert np.array_equal(x, a + b * 5.0)
    oracle_lib.lib().oracle_fmadd(P(x), P(b), P(c), 5.0, n)
    assert np.array_equal(x, b + c * 5.0)


def test_fmadd_is_fused():
    """a + b*c with one rounding: 1 + (1+eps)(1-eps) = 2 - eps^2 rounds to 2 when
    the product is rounded first, and to 2 - eps... with a single rounding."""
    a = np.array([-1.0])
    b = np.array([1.0 + EPS])
    x = a.copy()
    oracle_lib.lib().oracle_fmadd_inplace(P(x), P(b), 1.0 - EPS, 1)
    assert x[0] == -EPS * EPS          # exact: (1+e)(1-e) - 1 = -e^2


@pytest.mark.parametrize("k", [1, 3, 7])
def test_fmaddn_semantics(k):
    rng = np.random.default_rng(k)
    n = 64
    a, b = _ints(rng, n), _ints(rng, n)
    terms = [_ints(rng, n) for _ in range(k)]
    c = _ints(rng, k, 100)
    c[0] = 0
    c[k // 2] = 0
    x = a.copy()
    oracle_lib.lib().oracle_fmaddn_inplace(P(x), _arr(terms), P(c), k, n)
    assert np.array_equal(x, a + sum(t * ci for t, ci in zip(terms, c)))
    x = a.copy()
    oracle_lib.lib().oracle_fmaddn(P(x), P(b), _arr(terms), P(c), k, n)
    exp = a if np.all(c == 0) else b + sum(t * ci for t, ci in zip(terms, c))
    assert np.array_equal(x, exp)
    x = a.copy()
    oracle_lib.lib().oracle_fmaddn(P(x), None, _arr(terms), P(c), k, n)
    assert np.array_equal(x, sum(t * ci for t, ci in zip(terms, c)) + 0.0)


def test_fmaddn_all_zero_leaves_a():
    """nbody_engine::fmaddn with b != NULL and every d == 0 never writes a."""
    a = np.array([1.0, 2.0, 3.0])
    b = np.array([7.0, 7.0, 7.0])
    t = np.array([5.0, 5.0, 5.0])
    c = np.array([0.0])
    oracle_lib.lib().oracle_fmaddn(P(a), P(b), _arr([t]), P(c), 1, 3)
    assert list(a) == [1.0, 2.0, 3.0]


def test_fmaddn_corr_kat():
    n, k = 37, 10
    a = np.ones(n)
    corr = np.zeros(n)
    terms = [np.ones(n) for _ in range(k)]
    c = np.full(k, EPS / 2)
    oracle_lib.lib().oracle_fmaddn_corr(P(a), P(corr), _arr(terms), P(c), k, n)
    assert np.all(np.abs((1.0 + k * EPS / 2) - a) <= EPS)
    assert np.all(a > 1.0)
    plain = np.ones(n)
    oracle_lib.lib().oracle_fmaddn_inplace(P(plain), _arr(terms), P(c), k, n)
    assert np.all(plain == 1.0)                # uncompensated: every eps/2 is lost


def test_fmaxabs_and_clamp_kats():
    rng = np.random.default_rng(4)
    a = rng.integers(0, 10000, 100).astype(np.float64) - 9000
    assert oracle_lib.lib().oracle_fmaxabs(P(a), 100) == np.max(np.abs(a))
    n = 20
    y = np.where(np.arange(6 * n) & 1, 4.0, -4.0)
    oracle_lib.lib().oracle_clamp(P(y), n, 1.0)
    assert np.array_equal(y[:3 * n], np.where(np.arange(3 * n) & 1, 2.0, -2.0))
    assert np.array_equal(y[3 * n:], np.where(np.arange(3 * n, 6 * n) & 1, 4.0, -4.0))
