"""Barnes-Hut heap kd-tree on the GPU against the oracle restatement of
nbody_space_heap / nbody_space_heap_stackless / nbody_engine_simple_bh:
tree layout and node moments bit-exact, FP64 traversals bit-exact, engine
cross-checks as in test/engine/test_nbody_engine.cpp:1269-1321."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from hip_capi import Ctx
from nbody_amd import native, nbody_create_engine, nbody_data
from solver_cases import golden, max_abs_diff, run_case

pytestmark = pytest.mark.gpu

LAYOUT = {"heap": 1, "heap_stackless": 2}


@pytest.fixture(scope="module")
def ctx():
    c = Ctx(0)
    yield c
    c.close()


def _state(kind, n, seed):
    d = nbody_data()
    d.set_quiet(True)
    d.generate(kind, n, seed)
    return oracle_lib.state_from_data(d)


class GpuTree:
    def __init__(self, ctx, n):
        self.ctx, self.n = ctx, n
        self.h = C.c_void_p()
        native.check(ctx.lib.nbh_tree_create(ctx.h, n, C.byref(self.h)), "tree_create")

    def close(self):
        self.ctx.lib.nbh_tree_destroy(self.h)

    def build(self, dy, dm, ratio):
        native.check(self.ctx.lib.nbh_tree_build(self.ctx.h, self.h, dy.ptr, dm.ptr, ratio), "build")

    def refit(self, dy, ratio):
        native.check(self.ctx.lib.nbh_tree_refit(self.ctx.h, self.h, dy.ptr, ratio), "refit")

    def fcompute(self, dy, dm, layout, prec=0):
        df = self.ctx.alloc(6 * self.n)
        native.check(self.ctx.lib.nbh_bh_fcompute(self.ctx.h, self.h, dy.ptr, dm.ptr, df.ptr, layout, prec), "bh")
        return df.read()

    def arrays(self):
        s = 2 * self.n
        body = np.empty(s, dtype=np.int32)
        cm = np.empty(3 * s); mass = np.empty(s); crit = np.empty(s); bmin = np.empty(3 * s); bmax = np.empty(3 * s)
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        native.check(self.ctx.lib.nbh_tree_get(self.ctx.h, self.h, P(body), P(cm), P(mass), P(crit), P(bmin), P(bmax)))
        return dict(body_n=body.astype(np.int64), cm=cm.reshape(-1, 3), mass=mass, crit_r2=crit,
                    bmin=bmin.reshape(-1, 3), bmax=bmax.reshape(-1, 3))

    def counters(self):
        a, b = C.c_uint64(), C.c_uint64()
        native.check(self.ctx.lib.nbh_bh_counters(self.ctx.h, self.h, C.byref(a), C.byref(b)))
        return a.value, b.value


def _compare_trees(g, o, n):
    assert np.array_equal(g["body_n"][n:], o["body_n"][n:]), "leaf bodies differ"
    assert np.all(g["body_n"][1:n] == -1)
    for k in ("mass", "crit_r2"):
        assert np.array_equal(g[k][1:], o[k][1:]), k
    for k in ("cm", "bmin", "bmax"):
        assert np.array_equal(g[k][1:], o[k][1:]), k


@pytest.mark.parametrize("kind,n", [("ball", 1), ("ball", 2), ("ball", 16), ("plummer", 128), ("cube", 1024),
                                    ("two_plummer", 4096), ("plummer", 65536)])
def test_build_bit_exact(ctx, kind, n):
    y, m = _state(kind, n, 21)
    dy, dm = ctx.put(y), ctx.put(m)
    t = GpuTree(ctx, n)
    t.build(dy, dm, 3.1623)
    h = oracle_lib.Heap(n)
    h.build(y, m, 3.1623)
    _compare_trees(t.arrays(), h.arrays(), n)
    t.close()


def test_build_ties_follow_index_order(ctx):
    """Equal coordinates: both sides break ties by body index."""
    n = 256
    rng = np.random.default_rng(5)
    pos = rng.integers(0, 4, size=(n, 3)).astype(np.float64)   # heavy ties on every axis
    pos[:5, 0] = -0.0
    y = np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], np.zeros(3 * n)])
    m = rng.uniform(0.5, 1.5, n)
    dy, dm = ctx.put(y), ctx.put(m)
    t = GpuTree(ctx, n)
    t.build(dy, dm, 2.0)
    h = oracle_lib.Heap(n)
    h.build(y, m, 2.0)
    _compare_trees(t.arrays(), h.arrays(), n)
    t.close()


def test_refit_bit_exact(ctx):
    n = 4096
    y, m = _state("plummer", n, 3)
    dy, dm = ctx.put(y), ctx.put(m)
    t = GpuTree(ctx, n)
    t.build(dy, dm, 2.0)
    h = oracle_lib.Heap(n)
    h.build(y, m, 2.0)
    y2 = y.copy()
    y2[:3 * n] += np.random.default_rng(1).standard_normal(3 * n) * 1e-3    # bodies move, order kept
    dy.write(y2)
    t.refit(dy, 1.25)
    h.rebuild(y2, 1.25)
    _compare_trees(t.arrays(), h.arrays(), n)
    t.close()


@pytest.mark.parametrize("ratio", [1e8, 3.1623, 2.0, 1.25, 0.5])
@pytest.mark.parametrize("layout", ["heap", "heap_stackless"])
def test_traverse_fp64_bit_exact(ctx, ratio, layout):
    n = 2048
    y, m = _state("plummer", n, 8)
    dy, dm = ctx.put(y), ctx.put(m)
    t = GpuTree(ctx, n)
    t.build(dy, dm, ratio)
    f = t.fcompute(dy, dm, LAYOUT[layout])
    h = oracle_lib.Heap(n)
    h.build(y, m, ratio)
    f0 = h.fcompute(y, m, LAYOUT[layout], 1)
    assert np.array_equal(f, f0)
    inter, visits = t.counters()
    assert inter > 0 and visits >= inter
    if ratio == 1e8:
        assert inter == n * (n - 1)           # every leaf but the target itself
    t.close()


def test_traverse_fp32_close_to_fp64(ctx):
    n = 1 << 15
    y, m = _state("plummer", n, 4)
    dy, dm = ctx.put(y), ctx.put(m)
    t = GpuTree(ctx, n)
    t.build(dy, dm, 2.0)
    f64 = t.fcompute(dy, dm, 2, 0)
    f32 = t.fcompute(dy, dm, 2, 1)
    a64 = f64[3 * n:].reshape(3, n)
    a32 = f32[3 * n:].reshape(3, n)
    rel = np.linalg.norm(a32 - a64, axis=0) / np.linalg.norm(a64, axis=0)
    assert np.sqrt(np.mean(rel ** 2)) < 1e-5 and np.percentile(rel, 99) < 1e-5, rel.max()
    assert np.array_equal(f32[:3 * n], y[3 * n:])
    t.close()


def _sample_accel(ctx, dy, dm, n, idx):
    di = ctx.alloc((len(idx) + 1) // 2)
    raw = np.ascontiguousarray(idx, dtype=np.uint32)
    native.check(ctx.lib.nbh_memcpy_h2d(ctx.h, di.ptr, raw.ctypes.data_as(C.c_void_p), raw.nbytes))
    da = ctx.alloc(3 * len(idx))
    native.check(ctx.lib.nbh_direct_sample(ctx.h, dy.ptr, dm.ptr, n, di.ptr, len(idx), da.ptr), "sample")
    return da.read().reshape(-1, 3)


def test_direct_sample_bit_exact(ctx):
    n = 5000
    y, m = _state("plummer", n, 12)
    dy, dm = ctx.put(y), ctx.put(m)
    idx = np.random.default_rng(0).choice(n, 777, replace=False)
    got = _sample_accel(ctx, dy, dm, n, idx)
    assert np.array_equal(got, oracle_lib.accel_targets(y, m, idx))


def _rel(a, a0):
    return np.linalg.norm(a - a0, axis=1) / np.linalg.norm(a0, axis=1)


@pytest.mark.parametrize("ratio", [1.25, 2.0, 4.0])
def test_group_walk_at_least_as_accurate(ctx, ratio):
    """The group walk takes a refinement of each body's cells: its error vs the
    FP64 direct sum is no larger than the per-body walk's (up to FP32 noise),
    at the cost of more interactions."""
    n = 1 << 15
    y, m = _state("plummer", n, 14)
    dy, dm = ctx.put(y), ctx.put(m)
    t = GpuTree(ctx, n)
    t.build(dy, dm, ratio)
    f_body = t.fcompute(dy, dm, 2, 0)
    inter_body, _ = t.counters()
    df = ctx.alloc(6 * n)
    native.check(ctx.lib.nbh_bh_fcompute_group(ctx.h, t.h, dy.ptr, df.ptr), "group")
    f_group = df.read()
    inter_group, visits = t.counters()
    idx = np.random.default_rng(1).choice(n, 2048, replace=False)
    a0 = _sample_accel(ctx, dy, dm, n, idx)
    e_body = _rel(f_body[3 * n:].reshape(3, n).T[idx], a0)
    e_group = _rel(f_group[3 * n:].reshape(3, n).T[idx], a0)
    rms = lambda e: float(np.sqrt(np.mean(e ** 2)))
    assert rms(e_group) <= rms(e_body) * 1.01 + 2e-6, (rms(e_group), rms(e_body))
    assert np.percentile(e_group, 99) <= np.percentile(e_body, 99) * 1.01 + 2e-6
    assert inter_group >= inter_body
    assert np.array_equal(f_group[:3 * n], y[3 * n:])
    t.close()


def test_group_walk_ratio_1e8_is_direct(ctx):
    n = 4096
    y, m = _state("ball", n, 15)
    dy, dm = ctx.put(y), ctx.put(m)
    t = GpuTree(ctx, n)
    t.build(dy, dm, 1e8)
    df = ctx.alloc(6 * n)
    native.check(ctx.lib.nbh_bh_fcompute_group(ctx.h, t.h, dy.ptr, df.ptr), "group")
    inter, _ = t.counters()
    assert inter == n * n                 # every leaf, own group included (self term is 0)
    a = df.read()[3 * n:].reshape(3, n).T
    a0 = oracle_lib.fcompute_simple(y, m)[3 * n:].reshape(3, n).T
    e = _rel(a, a0)
    assert np.sqrt(np.mean(e ** 2)) < 1e-5 and e.max() < 1e-3
    t.close()


@pytest.mark.parametrize("rate", [0, 2])
@pytest.mark.parametrize("layout", ["heap", "heap_stackless"])
@pytest.mark.parametrize("ratio", [3.1623, 1e8])
def test_engine_vs_simple_bh(rate, layout, ratio):
    """test_nbody_engine.cpp:1287-1321: 128 bodies (make_universe 64), 2 steps,
    y scaled by 1.99 between them, rate 0 and 2 (refit on odd steps)."""
    d = nbody_data()
    d.set_quiet(True)
    d.make_universe(64, 100, 100, 100)
    e1 = oracle_lib.simple_bh_engine(ratio, "heap", "nested_tree", rate, 1)
    e2 = nbody_create_engine({"engine": "hip_bh", "distance_to_node_radius_ratio": ratio, "tree_build_rate": rate,
                              "traverse_type": "nested_tree", "tree_layout": layout})
    assert e1.init(d) and e2.init(d)
    n = e1.problem_size()
    y1, y2 = e1.create_buffer(8 * n), e2.create_buffer(8 * n)
    e1.copy_buffer(y1, e1.get_y())
    e2.copy_buffer(y2, e2.get_y())
    for step in range(2):
        outs = []
        for e, y, fill in ((e1, y1, 1e10), (e2, y2, -1e10)):
            f = e.create_buffer(8 * n)
            e.fill_buffer(f, fill)
            e.set_step(step)
            e.fcompute(0.0, y, f)
            outs.append(e.read_buffer(f))
            e.free_buffer(f)
        assert np.max(np.abs(outs[0] - outs[1])) <= 1e-13
        assert np.array_equal(outs[0], outs[1])
        e1.fmadd_inplace(y1, y1, 0.99)
        e2.fmadd_inplace(y2, y2, 0.99)
    assert not e2.has_error()


def test_ratio_1e8_equals_direct():
    """lambda = 1e8 opens every internal node: BH = direct sum (test_nbody_engine.cpp:1361-1396, 1e-11)."""
    d = nbody_data()
    d.set_quiet(True)
    d.make_universe(64, 100, 100, 100)
    e = nbody_create_engine({"engine": "hip_bh", "distance_to_node_radius_ratio": 1e8, "tree_layout": "heap_stackless"})
    assert e.init(d)
    y, m = oracle_lib.state_from_data(d)
    n = e.problem_size()
    ym, fm = e.create_buffer(8 * n), e.create_buffer(8 * n)
    e.write_buffer(ym, y)
    e.fcompute(0.0, ym, fm)
    assert np.max(np.abs(e.read_buffer(fm) - oracle_lib.fcompute_simple(y, m))) < 1e-11


def test_solver_golden_euler_bh():
    """The reference also runs the euler golden on its BH engine with lambda = 1e8
    (test/solver/test_nbody_solver.cpp:279-296)."""
    e = nbody_create_engine({"engine": "hip_bh", "distance_to_node_radius_ratio": 1e8})
    data, s = run_case(e, {"solver": "euler"})
    exp = golden("euler")
    assert exp.is_equal(data, 1e-12), max_abs_diff(exp, data)


def test_hip_bh_rejects_non_power_of_two():
    d = nbody_data()
    d.set_quiet(True)
    d.generate("ball", 100, 1)
    e = nbody_create_engine({"engine": "hip_bh"})
    assert e is not None
    assert not e.init(d)


@pytest.mark.slow
def test_two_plummer_1m_layout_bit_exact(ctx):
    """Two Plummer spheres of 2^19 bodies each: the GPU heap equals the oracle's."""
    n = 1 << 20
    y, m = _state("two_plummer", n, 3)
    dy, dm = ctx.put(y), ctx.put(m)
    t = GpuTree(ctx, n)
    t.build(dy, dm, 2.0)
    h = oracle_lib.Heap(n)
    h.build(y, m, 2.0)
    _compare_trees(t.arrays(), h.arrays(), n)
    t.close()


def test_engine_quadrupole_walk_under_rk4():
    """The hip_bh engine with the quadrupole group walk (far cells per body, or
    through the local expansion) drives rk4 (tree rebuilt every force
    evaluation) and its forces are closer to the FP64 direct sum than the
    monopole walk's at the same opening ratio."""
    n = 1 << 14
    errs = {}
    for mp in ("monopole", "quadrupole", "local"):
        d = nbody_data()
        d.set_quiet(True)
        d.generate("plummer", n, 16)
        params = {"engine": "hip_bh", "force_precision": "fp32", "bh_walk": "group",
                  "bh_multipole": "monopole" if mp == "monopole" else "quadrupole",
                  "distance_to_node_radius_ratio": 1.6}
        if mp == "local":
            params["bh_far_field"] = "local"
        e = nbody_create_engine(params)
        assert e.init(d)
        from nbody_amd import nbody_create_solver
        s = nbody_create_solver({"solver": "rk4", "max_step": 1e-3})
        s.set_engine(e)
        for _ in range(3):
            s.advise(1e-3)
        assert not e.has_error() and e.get_compute_count() == 12
        e.get_data(d)
        y, m = oracle_lib.state_from_data(d)
        f = e.create_buffer(48 * n)
        e.fcompute(0.0, e.get_y(), f)
        a = e.read_buffer(f)[3 * n:].reshape(3, n).T
        idx = np.random.default_rng(2).choice(n, 512, replace=False)
        a0 = oracle_lib.accel_targets(y, m, idx.astype(np.uint64), 0)
        errs[mp] = np.sqrt(np.mean((np.linalg.norm(a[idx] - a0, axis=1) / np.linalg.norm(a0, axis=1)) ** 2))
        e.free_buffer(f)
        s.destroy()
        e.destroy()
    assert errs["quadrupole"] < 0.3 * errs["monopole"], errs
    assert errs["local"] < 0.3 * errs["monopole"], errs

