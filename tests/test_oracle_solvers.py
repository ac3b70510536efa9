"""The product's solvers (libnbody.so) driven by the oracle `simple` engine
against the reference's own goldens (test/solver/test_nbody_solver.cpp),
solver equivalences (test/solvers_equality) and the extrapolator closed form
(test/extrapolator).  CPU only: pins the oracle's force law and vector ops
and the solver restatement at once.
"""
import numpy as np
import pytest

import oracle_lib
from nbody_amd import api, nbody_create_solver, nbody_data
from solver_cases import BUTCHER, CASES, golden, max_abs_diff, run_case


@pytest.mark.parametrize("name,params", CASES, ids=[c[0] for c in CASES])
def test_solver_golden_on_oracle(name, params):
    e = oracle_lib.simple_engine(1)
    data, s = run_case(e, params)
    exp = golden(name)
    assert exp.is_equal(data, 1e-12), f"{name}: max |diff| = {max_abs_diff(exp, data):.3e}"
    s.destroy()
    e.destroy()


def test_solver_factory_rejects_invalid():
    assert nbody_create_solver({"solver": "invalid"}) is None
    assert nbody_create_solver({"solver": "adams", "starter_solver": "invalid"}) is None
    assert nbody_create_solver({}) is None
    assert nbody_create_solver({"solver": "rkdp", "error_threshold": "x"}) is None
    assert nbody_create_solver({"solver": "adams", "rank": 6}) is None


@pytest.mark.parametrize("name", sorted(BUTCHER))
def test_butcher_table_check(name):
    """test/solver/test_nbody_solver.cpp:87-144: sum(b1) = sum(b2) = 1 and
    sum_j a_ij = c_i, to 10 eps."""
    s = nbody_create_solver({"solver": name})
    a, b1, b2, c, implicit, embedded = s.butcher_table()
    eps = 10 * np.finfo(np.float64).eps
    assert abs(b1.sum() - 1) < eps
    assert abs(b2.sum() - 1) < eps
    n = len(c)
    for i in range(n):
        jmax = n if implicit else i
        row = a[i, :jmax]
        acc, corr = 0.0, 0.0
        for v in row:                      # summation_k, nbody/summation.h:7-14
            t = v - corr
            ns = acc + t
            corr = (ns - acc) - t
            acc = ns
        amax = 1.0 if jmax == 0 else np.max(np.abs(row))
        assert abs(acc - c[i]) / amax < eps, (name, i)


def _universe64():
    d = nbody_data()
    d.set_quiet(True)
    d.make_universe(64, 100, 100, 100)
    return d


def _y(e):
    return e.read_buffer(e.get_y())


@pytest.mark.parametrize("pair", ["rk4-vs-butcher", "euler-vs-adams1"])
def test_solvers_equality(pair):
    """test/solvers_equality/test_nbody_solvers_equality.cpp:85-121."""
    lib = api.native.host_lib()
    if pair == "rk4-vs-butcher":
        s1 = nbody_create_solver({"solver": "rk4"})
        s2 = api.nbody_solver(lib.nbs_solver_create_rk_butcher(b"rk4"))
    else:
        s1 = nbody_create_solver({"solver": "euler"})
        s2 = api.nbody_solver(lib.nbs_solver_create_adams(1, 0))
    d1, d2 = _universe64(), _universe64()
    e1, e2 = oracle_lib.simple_engine(1), oracle_lib.simple_engine(1)
    assert e1.init(d1) and e2.init(d2)
    for s, e in ((s1, e1), (s2, e2)):
        s.set_time_step(1e-3, 1e-2)
        s.set_engine(e)
    assert np.array_equal(_y(e1), _y(e2))
    for _ in range(7):
        s1.advise(1e-2)
        s2.advise(1e-2)
    diff = _y(e1) - _y(e2)
    assert np.sqrt(np.max(diff * diff)) < 1e-13
    assert np.sqrt(np.sum(diff * diff)) / diff.size < 1e-16


@pytest.mark.parametrize("kind", ["berrut", "neville"])
def test_extrapolator_closed_form(kind):
    """test/extrapolator/test_nbody_extrapolator.cpp:53-95: f(x)=1/(1+x^2)+1
    extrapolated to x -> inf; error estimates decrease and bound the error."""
    import ctypes as C
    lib = api.native.host_lib()
    d = nbody_data()
    d.set_quiet(True)
    lib.nbs_data_resize(d.handle, 1)
    e = oracle_lib.simple_engine(1)
    assert e.init(d)
    y = e.create_buffer(e.problem_size() * 8)
    sc = (C.c_size_t * 12)(2, 4, 6, 8, 12, 16, 24, 32, 48, 64, 96, 128)
    ex = lib.nbs_extrapolator_create(kind.encode(), e.handle, 1, sc, 12)
    assert ex
    func = lambda x: 1.0 / (1.0 + x * x) + 1.0
    expected = func(np.finfo(np.float64).max)
    last_est = np.finfo(np.float64).max
    last_abs = np.finfo(np.float64).max
    for level in range(12):
        e.write_buffer(y, np.full(e.problem_size(), func(float(sc[level]))))
        lib.nbs_extrapolator_update_table(ex, level, y.handle)
        if level < 2:
            continue
        err = lib.nbs_extrapolator_estimate_error(ex, level)
        assert last_est >= err
        last_est = err
        lib.nbs_extrapolator_extrapolate(ex, level, y.handle)
        got = e.read_buffer(y)[0]
        err_abs = abs(got - expected)
        assert err_abs < 10 * err
        assert last_abs >= err_abs
        last_abs = err_abs
    if kind == "neville":
        assert last_abs < 1e-9
    lib.nbs_extrapolator_destroy(ex)
    e.free_buffer(y)
    assert not lib.nbs_extrapolator_create(b"", None, 0, None, 0)
