"""FP64 vector kernels and conservation sums through the C ABI, against the
oracle: odd lengths, unaligned pointers, aliasing, long term lists, empty
inputs and the NaN rule of fmaxabs."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from hip_capi import Ctx, dbl_array, ptr_array
from nbody_amd import native, nbody_data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Ctx(0)
    yield c
    c.close()


def _oracle_fmaddn(init, terms, coef):
    n = len(terms[0]) if terms else len(init)
    a = np.zeros(n) if init is None else init.copy()
    arr = (oracle_lib.PD * max(1, len(terms)))(*[oracle_lib.p(t) for t in terms])
    cf = np.ascontiguousarray(coef, dtype=np.float64)
    oracle_lib.lib().oracle_fmaddn_inplace(oracle_lib.p(a), arr, oracle_lib.p(cf), len(terms), n)
    return a


@pytest.mark.parametrize("n", [1, 2, 3, 1001, 65537])
@pytest.mark.parametrize("k", [0, 1, 7, 35, 60])
def test_fmaddn_lengths_terms(ctx, n, k):
    rng = np.random.default_rng(n * 100 + k)
    init = rng.standard_normal(n)
    terms = [rng.standard_normal(n) for _ in range(k)]
    coef = rng.standard_normal(k)
    da, di = ctx.alloc(n), ctx.put(init)
    dts = [ctx.put(t) for t in terms]
    cf, pcf = dbl_array(coef)
    native.check(ctx.lib.nbh_fmaddn(ctx.h, da.ptr, di.ptr, ptr_array([t.ptr for t in dts]), pcf, k, n))
    assert np.array_equal(da.read(), _oracle_fmaddn(init, terms, coef))
    # init = NULL starts from 0
    native.check(ctx.lib.nbh_fmaddn(ctx.h, da.ptr, None, ptr_array([t.ptr for t in dts]), pcf, k, n))
    assert np.array_equal(da.read(), _oracle_fmaddn(np.zeros(n), terms, coef))


def test_fmaddn_unaligned_and_aliased(ctx):
    n = 999
    rng = np.random.default_rng(1)
    base = rng.standard_normal(n + 1)
    t = rng.standard_normal(n + 1)
    da, dt = ctx.put(base), ctx.put(t)
    cf, pcf = dbl_array([0.5, -2.0])
    # a = a[1:] + 0.5 t[1:] - 2 a[1:]   (8-byte offset pointers; a also a source)
    a1, t1 = da.at(1), dt.at(1)
    native.check(ctx.lib.nbh_fmaddn(ctx.h, a1, a1, ptr_array([t1, a1]), pcf, 2, n))
    exp = base[1:].copy()
    exp = _oracle_fmaddn(exp, [t[1:].copy(), base[1:].copy()], [0.5, -2.0])
    got = da.read()
    assert got[0] == base[0]
    assert np.array_equal(got[1:], exp)


def test_fmaddn_corr_long(ctx):
    n, k = 777, 50
    rng = np.random.default_rng(2)
    a0 = rng.standard_normal(n)
    c0 = rng.standard_normal(n) * 1e-17
    terms = [rng.standard_normal(n) for _ in range(k)]
    coef = rng.standard_normal(k) * 1e-3
    da, dc = ctx.put(a0), ctx.put(c0)
    dts = [ctx.put(t) for t in terms]
    cf, pcf = dbl_array(coef)
    native.check(ctx.lib.nbh_fmaddn_corr(ctx.h, da.ptr, dc.ptr, ptr_array([t.ptr for t in dts]), pcf, k, n))
    ea, ec = a0.copy(), c0.copy()
    arr = (oracle_lib.PD * k)(*[oracle_lib.p(t) for t in terms])
    oracle_lib.lib().oracle_fmaddn_corr(oracle_lib.p(ea), oracle_lib.p(ec), arr, oracle_lib.p(cf), k, n)
    assert np.array_equal(da.read(), ea)
    assert np.array_equal(dc.read(), ec)


def test_streaming_sizes_bit_exact(ctx):
    """Calls that stream more than the 256 MB last-level cache take the
    nontemporal kernels: same results as the oracle bit for bit (fmaddn with an
    odd length and k = 2, aliased init; fmaddn_corr with k = 3)."""
    n = 12_000_001                      # 8 n (k + 2) bytes > 256 MB for k = 2
    rng = np.random.default_rng(9)
    init = rng.standard_normal(n)
    terms = [rng.standard_normal(n) for _ in range(2)]
    coef = np.array([0.25, -3.0])
    da = ctx.put(init)
    dts = [ctx.put(t) for t in terms]
    cf, pcf = dbl_array(coef)
    native.check(ctx.lib.nbh_fmaddn(ctx.h, da.ptr, da.ptr, ptr_array([t.ptr for t in dts]), pcf, 2, n))
    assert np.array_equal(da.read(), _oracle_fmaddn(init, terms, coef))
    m, k = 9_000_000, 3                 # 8 m (k + 4) bytes > 256 MB
    a0 = rng.standard_normal(m)
    c0 = rng.standard_normal(m) * 1e-17
    terms = [rng.standard_normal(m) for _ in range(k)]
    coef = rng.standard_normal(k) * 1e-3
    da, dc = ctx.put(a0), ctx.put(c0)
    dts = [ctx.put(t) for t in terms]
    cf, pcf = dbl_array(coef)
    native.check(ctx.lib.nbh_fmaddn_corr(ctx.h, da.ptr, dc.ptr, ptr_array([t.ptr for t in dts]), pcf, k, m))
    ea, ec = a0.copy(), c0.copy()
    arr = (oracle_lib.PD * k)(*[oracle_lib.p(t) for t in terms])
    oracle_lib.lib().oracle_fmaddn_corr(oracle_lib.p(ea), oracle_lib.p(ec), arr, oracle_lib.p(cf), k, m)
    assert np.array_equal(da.read(), ea)
    assert np.array_equal(dc.read(), ec)


@pytest.mark.parametrize("n", [1, 5, 4096, 1000003])
def test_maxabs(ctx, n):
    rng = np.random.default_rng(n)
    a = rng.standard_normal(n) * 100
    a[rng.integers(0, n)] = -1e6
    d = ctx.put(a)
    r = C.c_double(0)
    native.check(ctx.lib.nbh_maxabs(ctx.h, d.ptr, n, C.byref(r)))
    assert r.value == oracle_lib.lib().oracle_fmaxabs(oracle_lib.p(a), n) == 1e6


def test_maxabs_nan_rules(ctx):
    a = np.array([1.0, np.nan, -3.0, 2.0])
    d = ctx.put(a)
    r = C.c_double(0)
    native.check(ctx.lib.nbh_maxabs(ctx.h, d.ptr, 4, C.byref(r)))
    assert r.value == 3.0                     # NaN never compares greater
    a[0] = np.nan
    d.write(a)
    native.check(ctx.lib.nbh_maxabs(ctx.h, d.ptr, 4, C.byref(r)))
    assert np.isnan(r.value)                  # result seeded with |a[0]|
    native.check(ctx.lib.nbh_maxabs(ctx.h, d.ptr, 0, C.byref(r)))
    assert r.value == 0.0


def test_fill_clamp(ctx):
    n = 1001
    d = ctx.alloc(6 * n)
    native.check(ctx.lib.nbh_fill(ctx.h, d.ptr, 3.5, 6 * n))
    assert np.all(d.read() == 3.5)
    y = np.random.default_rng(3).uniform(-5, 5, 6 * n)
    d.write(y)
    native.check(ctx.lib.nbh_clamp(ctx.h, d.ptr, n, 2.0))
    exp = y.copy()
    oracle_lib.lib().oracle_clamp(oracle_lib.p(exp), n, 2.0)
    assert np.array_equal(d.read(), exp)


@pytest.mark.parametrize("kind,n", [("ball", 1024), ("plummer", 2048), ("cube", 777)])
def test_stats_vs_oracle(ctx, kind, n):
    data = nbody_data()
    data.set_quiet(True)
    data.generate(kind, n, 9)
    y, m = oracle_lib.state_from_data(data)
    pos, vel, _, _ = data.bodies()
    dy, dm = ctx.put(y), ctx.put(m)
    out = np.zeros(12)
    native.check(ctx.lib.nbh_stats(ctx.h, dy.ptr, dm.ptr, n, 1, out.ctypes.data_as(C.POINTER(C.c_double))))
    ref = oracle_lib.stats(pos, vel, m, True)
    scale = np.array([np.sum(m * np.linalg.norm(vel, axis=1))] * 3 +
                     [np.sum(m * np.linalg.norm(np.cross(pos, vel), axis=1))] * 3 +
                     [abs(ref[6]), abs(ref[7])] + [np.max(np.abs(pos))] * 3 + [1.0])
    scale[scale == 0] = 1.0
    assert np.all(np.abs(out - ref) / scale < 1e-14), (out, ref)
