"""Pins the oracle's force law and direct sum (nbody/nbody_data.cpp:35-55,
nbody/nbody_engine_simple.cpp:116-173) with closed forms, symmetry and a
long-double brute force -- checks that a dropped term, a sign or an index
mistake would fail."""
import math

import numpy as np
import pytest

import oracle_lib


def _y(pos, vel=None):
    pos = np.asarray(pos, dtype=np.float64)
    n = len(pos)
    vel = np.zeros((n, 3)) if vel is None else np.asarray(vel, dtype=np.float64)
    return np.ascontiguousarray(np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2]]))


def test_two_body_closed_form():
    m = np.array([3.0, 5.0])
    y = _y([[0.0, 0.0, 0.0], [2.0, 0.0, 0.0]], [[1, 2, 3], [4, 5, 6]])
    f = oracle_lib.fcompute_simple(y, m)
    n = 2
    assert np.array_equal(f[:3 * n], y[3 * n:])                 # f = [v ; a]
    a = f[3 * n:].reshape(3, n)
    assert a[0, 0] == pytest.approx(5.0 / 4.0, rel=1e-15)        # toward body 2, m2 / r^2
    assert a[0, 1] == pytest.approx(-3.0 / 4.0, rel=1e-15)
    assert np.all(a[1:] == 0.0)


def test_force_function_sign_and_clamp():
    out = np.empty(3)
    lib = oracle_lib.lib()
    v1 = np.array([1.0, 0.0, 0.0]); v2 = np.array([0.0, 0.0, 0.0])
    lib.oracle_force(oracle_lib.p(v1), oracle_lib.p(v2), 2.0, 3.0, oracle_lib.p(out))
    assert out[0] == pytest.approx(-6.0) and out[1] == 0 and out[2] == 0   # attraction: F on 1 points to 2
    v1 = np.array([1e-5, 0.0, 0.0])                                       # r^2 = 1e-10 < 1e-8: clamped
    lib.oracle_force(oracle_lib.p(v1), oracle_lib.p(v2), 1.0, 1.0, oracle_lib.p(out))
    assert out[0] == pytest.approx(-1e-5 / (1e-8 * math.sqrt(1e-8)), rel=1e-15)
    assert lib.oracle_potential_energy(oracle_lib.p(v1), oracle_lib.p(v2), 1.0, 1.0) == 0.0
    v1 = np.array([0.0, 3.0, 4.0])
    assert lib.oracle_potential_energy(oracle_lib.p(v1), oracle_lib.p(v2), 2.0, 5.0) == pytest.approx(-2.0)


def test_symmetric_configuration():
    """Six equal masses on the axes around a central body: the centre feels no force,
    each outer body is pulled radially inward."""
    pos = [[0, 0, 0], [1, 0, 0], [-1, 0, 0], [0, 1, 0], [0, -1, 0], [0, 0, 1], [0, 0, -1]]
    m = np.full(7, 2.0)
    f = oracle_lib.fcompute_simple(_y(pos), m)
    a = f[21:].reshape(3, 7)
    assert np.allclose(a[:, 0], 0.0, atol=1e-15)
    for i in range(1, 7):
        r = np.array(pos[i], dtype=float)
        ai = a[:, i]
        assert np.linalg.norm(np.cross(ai, r)) < 1e-14 and np.dot(ai, r) < 0


def test_newton_third_law():
    rng = np.random.default_rng(1)
    n = 200
    pos = rng.standard_normal((n, 3))
    m = rng.uniform(0.1, 2.0, n)
    a = oracle_lib.fcompute_simple(_y(pos), m)[3 * n:].reshape(3, n)
    total = (a * m).sum(axis=1)
    assert np.linalg.norm(total) / np.sum(m * np.linalg.norm(a, axis=0)) < 1e-14


def test_long_double_brute_force():
    rng = np.random.default_rng(2)
    n = 64
    pos = rng.uniform(-1, 1, (n, 3))
    m = rng.uniform(0.5, 1.5, n)
    a = oracle_lib.fcompute_simple(_y(pos), m)[3 * n:].reshape(3, n)
    P = pos.astype(np.longdouble)
    M = m.astype(np.longdouble)
    for i in range(n):
        d = P - P[i]
        r2 = (d * d).sum(axis=1)
        r2[i] = 1
        w = M / (r2 * np.sqrt(r2))
        w[i] = 0
        ai = (d * w[:, None]).sum(axis=0)
        assert np.allclose(a[:, i].astype(np.longdouble), ai, rtol=1e-13, atol=0)


def test_openmp_equals_simple():
    rng = np.random.default_rng(3)
    n = 300
    y = _y(rng.standard_normal((n, 3)), rng.standard_normal((n, 3)))
    m = rng.uniform(0.1, 1, n)
    assert np.array_equal(oracle_lib.fcompute_simple(y, m, 1), oracle_lib.fcompute_simple(y, m, 4))
    idx = np.array([0, 17, 299])
    acc = oracle_lib.accel_targets(y, m, idx, 2)
    f = oracle_lib.fcompute_simple(y, m)
    assert np.array_equal(acc.T, f[3 * n:].reshape(3, n)[:, idx])
