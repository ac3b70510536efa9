"""The `hip` engine through the reference's engine test suite
(test/engine/test_nbody_engine.cpp:10-1011), plus bit-exact comparisons of
every FP64 op with the oracle `simple` engine on the same inputs.
"""
import numpy as np
import pytest

import oracle_lib
from nbody_amd import fake_memory, log_count, nbody_create_engine, nbody_data

pytestmark = pytest.mark.gpu

EPS = np.finfo(np.float64).eps
RNG = np.random.default_rng(12345)


def _rand_int(n, lo=0, hi=10000):
    return RNG.integers(lo, hi, size=n).astype(np.float64)


@pytest.fixture(scope="module")
def universe():
    d = nbody_data()
    d.set_quiet(True)
    d.make_universe(64, 100, 100, 100)       # 128 bodies, test_nbody_engine.cpp:684-690
    return d


@pytest.fixture(scope="module", params=["hip", "hip_bh_1e8"])
def engine(request, universe):
    if request.param == "hip":
        e = nbody_create_engine({"engine": "hip"})
    else:
        e = nbody_create_engine({"engine": "hip_bh", "distance_to_node_radius_ratio": 1e8,
                                 "tree_layout": "heap_stackless"})
    assert e is not None
    assert e.init(universe)
    yield e
    e.destroy()


def test_mem(engine):
    m = engine.create_buffer(1024)
    assert m is not None
    engine.free_buffer(m)


def test_memcpy(engine):
    data = np.arange(8, dtype=np.float64)
    mem = engine.create_buffer(64)
    sub = engine.create_buffer(32)
    engine.write_buffer(mem, data)
    assert np.array_equal(engine.read_buffer(mem), data)
    engine.free_buffer(mem)
    engine.free_buffer(sub)


def test_copy_buffer(engine):
    n = engine.problem_size()
    a, b = engine.create_buffer(8 * n), engine.create_buffer(8 * n)
    d = np.arange(n, dtype=np.float64)
    engine.write_buffer(a, d)
    engine.copy_buffer(b, a)
    assert np.array_equal(engine.read_buffer(b), d)
    engine.free_buffers([a, b])


def test_fill_buffer(engine):
    m = engine.create_buffer(33 * 8)
    engine.fill_buffer(m, 777.0)
    assert np.all(engine.read_buffer(m) == 777.0)
    engine.free_buffer(m)


def test_fmadd_inplace(engine):
    n = engine.problem_size()
    a, b = _rand_int(n), _rand_int(n)
    ma, mb = engine.create_buffer(8 * n), engine.create_buffer(8 * n)
    engine.write_buffer(ma, a)
    engine.write_buffer(mb, b)
    engine.fmadd_inplace(ma, mb, 5.0)
    assert np.all(np.abs((a + 5.0 * b) - engine.read_buffer(ma)) <= EPS)
    engine.free_buffers([ma, mb])


def test_fmadd(engine):
    n = engine.problem_size()
    a, b, c = _rand_int(n), _rand_int(n), _rand_int(n)
    ma, mb, mc = (engine.create_buffer(8 * n) for _ in range(3))
    for m, v in ((ma, a), (mb, b), (mc, c)):
        engine.write_buffer(m, v)
    engine.fmadd(ma, mb, mc, 5.0)
    assert np.all(np.abs((b + c * 5.0) - engine.read_buffer(ma)) <= EPS)
    engine.free_buffers([ma, mb, mc])


@pytest.mark.parametrize("csize", [1, 3, 7])
def test_fmaddn_inplace(engine, csize):
    """c[0] = c[csize/2] = 0 as in test_nbody_engine.cpp:188-245."""
    n = engine.problem_size()
    a = _rand_int(n)
    b = [_rand_int(n) for _ in range(csize)]
    c = _rand_int(csize, 0, 100)
    c[0] = 0
    c[csize // 2] = 0
    ma = engine.create_buffer(8 * n)
    mb = engine.create_buffers(8 * n, csize)
    engine.write_buffer(ma, a)
    for m, v in zip(mb, b):
        engine.write_buffer(m, v)
    engine.fmaddn_inplace(ma, mb, c, csize)
    exp = a.copy()
    for k in range(csize):
        exp += b[k] * c[k]
    assert np.all(np.abs(exp - engine.read_buffer(ma)) <= EPS)
    engine.free_buffer(ma)
    engine.free_buffers(mb)


@pytest.mark.parametrize("csize", [1, 3, 7])
@pytest.mark.parametrize("null_b", [False, True])
def test_fmaddn(engine, csize, null_b):
    n = engine.problem_size()
    a, b = _rand_int(n), _rand_int(n)
    c = [_rand_int(n) for _ in range(csize)]
    d = _rand_int(csize, 0, 100)
    d[0] = 0
    d[csize // 2] = 0
    ma, mb = engine.create_buffer(8 * n), engine.create_buffer(8 * n)
    mc = engine.create_buffers(8 * n, csize)
    engine.write_buffer(ma, a)
    engine.write_buffer(mb, b)
    for m, v in zip(mc, c):
        engine.write_buffer(m, v)
    engine.fmaddn(ma, None if null_b else mb, mc, d, csize)
    exp = np.zeros(n) if null_b else b.copy()
    for k in range(csize):
        exp += c[k] * d[k]
    got = engine.read_buffer(ma)
    if not null_b and csize == 1:
        exp = a          # only term has d = 0: a is left as it was (nbody_engine.cpp:94-107)
    assert np.all(np.abs(exp - got) <= EPS)
    engine.free_buffers([ma, mb])
    engine.free_buffers(mc)


def test_fmaddn_corr(engine):
    """test_nbody_engine.cpp:377-437: ten terms of eps/2 added to 1 need compensation."""
    n = engine.problem_size()
    csize = 10
    ma, mcorr = engine.create_buffer(8 * n), engine.create_buffer(8 * n)
    mb = engine.create_buffers(8 * n, csize)
    engine.write_buffer(ma, np.ones(n))
    engine.fill_buffer(mcorr, 0.0)
    for m in mb:
        engine.write_buffer(m, np.ones(n))
    c = np.full(csize, EPS / 2)
    engine.fmaddn_corr(ma, mcorr, mb, c, csize)
    got = engine.read_buffer(ma)
    assert np.all(np.abs((1.0 + csize * EPS / 2) - got) <= EPS)
    assert np.all(got > 1.0)
    engine.free_buffers([ma, mcorr])
    engine.free_buffers(mb)


def test_fmaxabs(engine):
    n = engine.problem_size()
    a = _rand_int(n) - 9000
    m = engine.create_buffer(8 * n)
    engine.write_buffer(m, a)
    assert abs(engine.fmaxabs(m, 2878767678687.0) - np.max(np.abs(a))) <= EPS
    engine.free_buffer(m)


def test_clamp(engine):
    n = engine.problem_size()
    a = np.where(np.arange(n) & 1, 4.0, -4.0)
    m = engine.create_buffer(8 * n)
    engine.write_buffer(m, a)
    engine.clamp(m, 1.0)
    b = engine.read_buffer(m)
    half = n // 2
    assert np.array_equal(b[:half], np.where(np.arange(half) & 1, 2.0, -2.0))
    assert np.array_equal(b[half:], a[half:])
    engine.free_buffer(m)


def _fcompute(e, y):
    n = e.problem_size()
    ym, fm = e.create_buffer(8 * n), e.create_buffer(8 * n)
    e.write_buffer(ym, y)
    e.fill_buffer(fm, -1e10)
    e.fcompute(0.0, ym, fm)
    f = e.read_buffer(fm)
    e.free_buffers([ym, fm])
    return f


def test_fcompute_vs_simple(engine, universe):
    """test_nbody_engine.cpp:472-558: against `simple` at 1e-13, two rounds with
    y scaled by 1.99 in between; the FP64 direct kernel is bit-exact."""
    y, m = oracle_lib.state_from_data(universe)
    # BH with lambda = 1e8 sums the same pairs in tree order: 1e-11 as in the reference (:1361-1396)
    eps = 1e-13 if engine.type_name() == "nbody_engine_hip" else 1e-11
    for step in range(2):
        engine.set_step(step)
        f = _fcompute(engine, y)
        f0 = oracle_lib.fcompute_simple(y, m)
        assert np.max(np.abs(f - f0)) <= eps
        if engine.type_name() == "nbody_engine_hip":
            assert np.array_equal(f, f0), "FP64 direct sum must equal the CPU order bit for bit"
        y = y + y * 0.99


def test_negative_branches(engine):
    """test_nbody_engine.cpp:761-1011: misuse is logged and returns, nothing crashes."""
    ps = engine.problem_size()
    fake0 = fake_memory(0)
    before = log_count()
    real = engine.create_buffer(ps)
    reals = engine.create_buffers(ps, 1)
    engine.fcompute(0, fake0, real)
    engine.fcompute(0, real, fake0)
    engine.read_buffer(fake0, np.empty(1))
    engine.write_buffer(fake0, None)
    engine.copy_buffer(fake0, real)
    engine.copy_buffer(real, fake0)
    engine.fmadd_inplace(fake0, real, 1)
    engine.fmadd_inplace(real, fake0, 1)
    engine.fmadd(fake0, real, real, 0)
    engine.fmadd(real, real, fake0, 0)
    engine.fmadd(real, fake0, real, 0)
    engine.fmaddn_inplace(fake0, reals, [0.0], 1)
    engine.fmaddn_inplace(real, reals, None, 1)
    engine.fmaddn_inplace(real, [fake0], [0.0], 1)
    engine.fmaddn_inplace(real, reals, [0.0], 100)
    engine.fmaddn_inplace(real, [None], [1.0], 1)
    engine.fmaddn_corr(None, real, reals, [0.0], 1)
    engine.fmaddn_corr(real, None, reals, [0.0], 1)
    engine.fmaddn_corr(real, real, [None], [0.0], 1)
    engine.fmaddn_corr(real, real, reals, None, 1)
    engine.fmaddn_corr(real, real, reals, [0.0], 100)
    engine.fmaddn_corr(real, real, [None], [1.0], 1)
    engine.fmaddn(fake0, real, reals, [0.0], 200000)
    engine.fmaddn(fake0, real, reals, [0.0], 1)
    engine.fmaddn(fake0, None, reals, [0.0], 1)
    engine.fmaddn(real, real, reals, None, 1)
    engine.fmaddn(real, None, reals, None, 1)
    engine.fmaddn(real, real, [fake0], [0.0], 1)
    engine.fmaddn(real, None, [fake0], [0.0], 1)
    engine.fmaddn(real, fake0, reals, [0.0], 1)
    engine.fmaddn(real, real, [None], [1.0], 1)
    engine.fmaxabs(fake0)
    assert log_count() > before
    assert not engine.has_error()
    engine.free_buffer(real)
    engine.free_buffers(reals)


def test_vector_ops_bit_exact_vs_oracle(engine):
    """Fused multi-term kernels equal the oracle's one-term-at-a-time decomposition."""
    n = engine.problem_size()
    rng = np.random.default_rng(7)
    a = rng.standard_normal(n)
    b = rng.standard_normal(n)
    terms = [rng.standard_normal(n) for _ in range(9)]
    coef = rng.standard_normal(9)
    coef[[2, 5]] = 0.0
    ma, mb = engine.create_buffer(8 * n), engine.create_buffer(8 * n)
    mt = engine.create_buffers(8 * n, 9)
    for m, v in zip(mt, terms):
        engine.write_buffer(m, v)
    lib = oracle_lib.lib()
    import ctypes as C
    arr = (oracle_lib.PD * 9)(*[oracle_lib.p(t) for t in terms])
    # fmaddn with b
    engine.write_buffer(ma, a)
    engine.write_buffer(mb, b)
    engine.fmaddn(ma, mb, mt, coef, 9)
    exp = a.copy()
    lib.oracle_fmaddn(oracle_lib.p(exp), oracle_lib.p(b), arr, oracle_lib.p(coef), 9, n)
    assert np.array_equal(engine.read_buffer(ma), exp)
    # fmaddn with b = NULL
    engine.fmaddn(ma, None, mt, coef, 9)
    lib.oracle_fmaddn(oracle_lib.p(exp), None, arr, oracle_lib.p(coef), 9, n)
    assert np.array_equal(engine.read_buffer(ma), exp)
    # fmaddn_inplace
    engine.write_buffer(ma, a)
    engine.fmaddn_inplace(ma, mt, coef, 9)
    exp = a.copy()
    lib.oracle_fmaddn_inplace(oracle_lib.p(exp), arr, oracle_lib.p(coef), 9, n)
    assert np.array_equal(engine.read_buffer(ma), exp)
    # fmaddn_corr
    corr0 = rng.standard_normal(n) * 1e-17
    engine.write_buffer(ma, a)
    engine.write_buffer(mb, corr0)
    engine.fmaddn_corr(ma, mb, mt, coef, 9)
    exp, ecorr = a.copy(), corr0.copy()
    lib.oracle_fmaddn_corr(oracle_lib.p(exp), oracle_lib.p(ecorr), arr, oracle_lib.p(coef), 9, n)
    assert np.array_equal(engine.read_buffer(ma), exp)
    assert np.array_equal(engine.read_buffer(mb), ecorr)
    engine.free_buffers([ma, mb])
    engine.free_buffers(mt)


def test_factory_rejects_invalid():
    assert nbody_create_engine({"engine": "invalid"}) is None
    assert nbody_create_engine({"engine": "hip", "device": "0,0"}) is None
    assert nbody_create_engine({"engine": "hip", "device": "99"}) is None
    assert nbody_create_engine({"engine": "hip", "force_precision": "fp16"}) is None
    assert nbody_create_engine({"engine": "hip_bh", "tree_layout": "tree"}) is None
    assert nbody_create_engine({"engine": "hip_bh", "traverse_type": "x"}) is None
    assert nbody_create_engine({"engine": "hip_bh", "distance_to_node_radius_ratio": "abc"}) is None
