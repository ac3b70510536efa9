"""The Morton radix tree on the GPU (nbh_tree_build_radix) and the group walks on it.

The build is held bit for bit to the oracle's restatement (oracle_radix_build,
itself pinned by its definition in test_oracle_radix.py): sorted body order,
children, parents, masses, mass centres, boxes and opening radii.  The walks
are held to the FP64 direct sum: with an opening ratio of 1e8 every body is
taken one by one (a coverage check: the FP32 walk is then a direct sum), and at
the ratios of use the error is bounded and, for a centrally concentrated state,
below that of the same walk on the heap kd-tree."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from hip_capi import Ctx
from nbody_amd import native, nbody_create_engine, nbody_create_solver, nbody_data

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    c = Ctx(0)
    yield c
    c.close()


def _state(kind, n, seed):
    if kind in ("plummer", "ball", "cube"):
        d = nbody_data()
        d.set_quiet(True)
        d.generate(kind, n, seed)
        return oracle_lib.state_from_data(d)
    rng = np.random.default_rng(seed)
    pos = rng.standard_normal((5, 3))[rng.integers(0, 5, n)]      # coincident bodies
    y = np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], np.zeros(3 * n)])
    return y, rng.random(n) + 0.1


class RTree:
    def __init__(self, ctx, y, m, ratio, radix=True, lean=False):
        self.ctx, self.n = ctx, len(m)
        self.h = C.c_void_p()
        native.check(ctx.lib.nbh_tree_create(ctx.h, self.n, C.byref(self.h)), "tree_create")
        if lean:
            native.check(ctx.lib.nbh_tree_set_radix_lean(ctx.h, self.h, 1), "lean")
        self.dy, self.dm = ctx.put(y), ctx.put(m)
        build = ctx.lib.nbh_tree_build_radix if radix else ctx.lib.nbh_tree_build
        native.check(build(ctx.h, self.h, self.dy.ptr, self.dm.ptr, ratio), "build")

    def close(self):
        self.ctx.lib.nbh_tree_destroy(self.h)

    def arrays(self):
        s = 2 * self.n
        body = np.empty(s, np.int32)
        cm = np.empty(3 * s); mass = np.empty(s); crit = np.empty(s); bmin = np.empty(3 * s); bmax = np.empty(3 * s)
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        native.check(self.ctx.lib.nbh_tree_get(self.ctx.h, self.h, P(body), P(cm), P(mass), P(crit), P(bmin),
                                               P(bmax)), "tree_get")
        child = np.empty(2 * self.n, np.int32)
        parent = np.empty(2 * self.n, np.int32)
        native.check(self.ctx.lib.nbh_tree_get_children(self.ctx.h, self.h, P(child), P(parent)), "children")
        return dict(body=body, cm=cm.reshape(-1, 3), mass=mass, crit=crit, bmin=bmin.reshape(-1, 3),
                    bmax=bmax.reshape(-1, 3), child=child.reshape(-1, 2), parent=parent)

    def walk(self, kind, near_k=3.0):
        df = self.ctx.alloc(6 * self.n)
        L, h, y = self.ctx.lib, self.ctx.h, self.dy.ptr
        if kind == "group":
            st = L.nbh_bh_fcompute_group(h, self.h, y, df.ptr)
        elif kind == "quad":
            st = L.nbh_bh_fcompute_group_quad(h, self.h, y, df.ptr, near_k)
        elif kind == "local":
            st = L.nbh_bh_fcompute_group_local(h, self.h, y, df.ptr, near_k)
        else:
            st = L.nbh_bh_fcompute_group_local2(h, self.h, y, df.ptr, near_k)
        native.check(st, kind)
        return df.read()

    def counters(self):
        c = (C.c_uint64 * 5)()
        native.check(self.ctx.lib.nbh_bh_walk_counters(self.ctx.h, self.h, c))
        return list(c)


@pytest.mark.parametrize("kind,n,seed", [("plummer", 2, 7), ("plummer", 3, 8), ("plummer", 7, 9), ("plummer", 100, 10),
                                         ("plummer", 64, 1), ("plummer", 3000, 2), ("plummer", 4096, 3),
                                         ("ball", 1 << 15, 4), ("plummer", 1 << 15, 5), ("coincident", 2048, 6)])
def test_radix_build_bit_exact(ctx, kind, n, seed):
    y, m = _state(kind, n, seed)
    ratio = 1.6
    t = RTree(ctx, y, m, ratio)
    g = t.arrays()
    o = oracle_lib.radix_build(y, m, ratio)
    ids = slice(1, 2 * n)
    assert np.array_equal(g["body"][n:], o["order"].astype(np.int32))
    assert np.all(g["body"][1:n] == -1)
    assert np.array_equal(g["child"][1:n], o["child"][1:n])
    assert np.array_equal(g["parent"][2:], o["parent"][2:])
    for k in ("cm", "mass", "crit", "bmin", "bmax"):
        assert np.array_equal(g[k][ids], o[k][ids]), k
    t.close()


@pytest.mark.parametrize("kind,n,seed", [("plummer", 3, 8), ("plummer", 100, 10), ("plummer", 3000, 2),
                                         ("plummer", 1 << 15, 5), ("coincident", 2500, 6)])
def test_radix_lean_walks_equal(ctx, kind, n, seed):
    """A lean build (FP64 node arrays only where the build reads them back) gives
    the group walks the same records: every walk equals the full build's bit for
    bit, after the build and after a refit of moved bodies, and the root's
    moments agree."""
    y, m = _state(kind, n, seed)
    ratio = 1.6
    full, lean = RTree(ctx, y, m, ratio), RTree(ctx, y, m, ratio, lean=True)
    y2 = y.copy()
    y2[:3 * n] += np.random.default_rng(3).standard_normal(3 * n) * 1e-3
    for step in range(2):
        for walk in ("group", "quad", "local", "local2"):
            assert np.array_equal(full.walk(walk, 2.0), lean.walk(walk, 2.0)), (step, walk)
        a, b = full.arrays(), lean.arrays()
        for k in ("cm", "mass", "crit", "bmin", "bmax"):
            assert np.array_equal(a[k][1], b[k][1]), k
        assert np.array_equal(a["body"], b["body"])
        for t in (full, lean):
            t.dy.write(y2)
            native.check(ctx.lib.nbh_tree_refit(ctx.h, t.h, t.dy.ptr, ratio), "refit")
    full.close()
    lean.close()


def test_radix_refit_keeps_structure(ctx):
    """nbh_tree_refit on a Morton tree recomputes the moments for moved bodies
    on the same structure: equal to a build when nothing moved, and sums over
    each node's bodies after a move."""
    n = 4096
    y, m = _state("plummer", n, 11)
    ratio = 1.6
    t = RTree(ctx, y, m, ratio)
    before = t.arrays()
    native.check(ctx.lib.nbh_tree_refit(ctx.h, t.h, t.dy.ptr, ratio), "refit")
    same = t.arrays()
    for k in ("cm", "mass", "crit", "bmin", "bmax", "child", "body"):
        assert np.array_equal(before[k], same[k]), k
    y2 = y.copy()
    y2[:3 * n] += np.random.default_rng(1).standard_normal(3 * n) * 1e-3
    t.dy.write(y2)
    native.check(ctx.lib.nbh_tree_refit(ctx.h, t.h, t.dy.ptr, ratio), "refit")
    after = t.arrays()
    assert np.array_equal(after["child"], before["child"]) and np.array_equal(after["body"], before["body"])
    pos = np.stack([y2[:n], y2[n:2 * n], y2[2 * n:3 * n]], 1)[after["body"][n:]]
    # leaf ranges of every internal id from the structure
    first = np.zeros(2 * n, np.int64); last = np.zeros(2 * n, np.int64)
    first[n:] = last[n:] = np.arange(n)
    order, stack = [], [1]
    while stack:
        k = stack.pop(); order.append(k)
        stack += [int(c) for c in after["child"][k] if c < n]
    for k in reversed(order):
        l, r = after["child"][k]
        first[k], last[k] = first[l], last[r]
    for k in order[::97]:
        p = pos[first[k]:last[k] + 1]
        assert np.array_equal(after["bmin"][k], p.min(0)) and np.array_equal(after["bmax"][k], p.max(0))
    t.close()


def _direct(y, m, idx):
    return oracle_lib.accel_targets(y, m, idx.astype(np.uint64), 8)


def _rms(f, a0, n, idx):
    a = f[3 * n:].reshape(3, n).T[idx]
    e = np.linalg.norm(a - a0, axis=1) / np.linalg.norm(a0, axis=1)
    return float(np.sqrt(np.mean(e ** 2))), float(np.percentile(e, 99))


@pytest.mark.parametrize("walk", ["group", "quad", "local", "local2"])
@pytest.mark.parametrize("kind,n", [("plummer", 5), ("plummer", 100), ("plummer", 2048), ("plummer", 3001),
                                    ("coincident", 2500)])
def test_radix_walk_coverage(ctx, walk, kind, n):
    """Opening ratio 1e8: every cell is opened down to its bodies, so each walk
    is an FP32 direct sum (every body counted once), within the direct FP32
    kernel's bound; also with walk groups smaller than 64 bodies (n not a
    multiple of 64, bodies on a few points)."""
    y, m = _state(kind, n, 21)
    t = RTree(ctx, y, m, 1e8)
    idx = np.arange(0, n, 7)
    e = _rms(t.walk(walk), _direct(y, m, idx), n, idx)
    assert e[0] < 1e-5 and e[1] < 1e-5, e
    assert t.counters()[2] == 0
    t.close()


@pytest.mark.parametrize("walk,theta,bound", [("group", 0.5, 8e-3), ("quad", 0.6, 2.5e-3), ("local", 0.6, 2.5e-3),
                                              ("local2", 0.6, 2.5e-3)])
def test_radix_walk_accuracy(ctx, walk, theta, bound):
    """Plummer sphere, 2^16 bodies: sampled accelerations against the FP64 direct
    sum, beside the same walk on the heap kd-tree, whose count-balanced cells
    span from the core to the outskirts: at the same opening angle the Morton
    tree's compact cells need a third of the interactions or fewer, for an error
    within a quarter of the heap tree's."""
    n = 1 << 16
    y, m = _state("plummer", n, 22)
    idx = np.sort(np.random.default_rng(5).choice(n, 2048, replace=False))
    a0 = _direct(y, m, idx)
    t = RTree(ctx, y, m, 1.0 / theta)
    er = _rms(t.walk(walk), a0, n, idx)
    cr = t.counters()
    t.close()
    h = RTree(ctx, y, m, 1.0 / theta, radix=False)
    eh = _rms(h.walk(walk), a0, n, idx)
    ch = h.counters()
    h.close()
    print(f"\n[morton tree] {walk} theta {theta}: rms/p99 {er[0]:.3e}/{er[1]:.3e}, interactions per body "
          f"{cr[0] / n:.0f}, spills {cr[2]}; heap tree {eh[0]:.3e}/{eh[1]:.3e}, {ch[0] / n:.0f} per body")
    assert er[0] < bound and cr[2] == 0
    assert er[0] < 1.25 * eh[0]
    assert 3 * cr[0] < ch[0]


def test_radix_engine_key():
    """hip_bh with bh_tree=morton: rk4 steps run on the Morton tree and follow the
    heap tree's trajectory to the walks' accuracy; bad combinations are refused."""
    n = 8192
    d = nbody_data()
    d.set_quiet(True)
    d.generate("plummer", n, 31)
    base = {"engine": "hip_bh", "force_precision": "fp32", "bh_walk": "group", "distance_to_node_radius_ratio": 2.0,
            "bh_multipole": "quadrupole", "bh_far_field": "local"}
    assert nbody_create_engine({**base, "bh_tree": "octree"}) is None
    assert nbody_create_engine({"engine": "hip_bh", "bh_tree": "morton"}) is None      # needs the group walk
    out = []
    for tree in ("heap", "morton"):
        dd = nbody_data()
        dd.set_quiet(True)
        dd.generate("plummer", n, 31)
        e = nbody_create_engine({**base, "bh_tree": tree})
        assert e is not None and e.init(dd)
        s = nbody_create_solver({"solver": "rk4", "max_step": 1e-3})
        s.set_engine(e)
        for _ in range(3):
            s.advise(1e-3)
        assert not e.has_error()
        e.get_data(dd)
        out.append(dd.bodies()[0].copy())
        s.destroy()
        e.destroy()
    dev = np.linalg.norm(out[0] - out[1], axis=1)
    assert np.percentile(dev, 99) < 1e-7 and dev.max() < 1e-5, (np.percentile(dev, 99), dev.max())
    # any body count: the Morton tree's walk groups are subtrees of up to 64 bodies
    dd = nbody_data()
    dd.set_quiet(True)
    dd.generate("plummer", 3000, 32)
    e = nbody_create_engine({**base, "bh_tree": "morton"})
    assert e is not None and e.init(dd)
    s = nbody_create_solver({"solver": "rk4", "max_step": 1e-3})
    s.set_engine(e)
    s.advise(1e-3)
    assert not e.has_error()
    e.get_data(dd)
    assert np.all(np.isfinite(dd.bodies()[0]))
    s.destroy()
    e.destroy()


def test_radix_two_level_stack_overflow_falls_back(ctx, monkeypatch):
    """The Morton tree's super-group walk run out of stack (forced through the
    library's NBH_TEST_SUPER_STACK hook, 64 entries): the walk groups of such a
    super-group ignore its partial list and expansion and walk one level, so
    every body gets either the one-level local walk's result or the normal
    two-level result, bit for bit, some of them the former, and the fallbacks are
    counted in walk counter out[2]."""
    n, near_k = 1 << 17, 3.0
    y, m = _state("plummer", n, 46)
    t = RTree(ctx, y, m, 1.0 / 0.3)         # a small opening angle: the super-group walks open many nodes
    one = t.walk("local", near_k)
    monkeypatch.setenv("NBH_TEST_SUPER_STACK", "64")
    two = t.walk("local2", near_k)
    c = t.counters()
    monkeypatch.delenv("NBH_TEST_SUPER_STACK")
    normal = t.walk("local2", near_k)
    assert t.counters()[2] == 0 and not np.array_equal(normal, one)
    a1, a2, an = (v[3 * n:].reshape(3, n) for v in (one, two, normal))
    same_one = np.all(a2 == a1, axis=0)
    same_normal = np.all(a2 == an, axis=0)
    assert np.all(same_one | same_normal)
    fell = int(np.sum(same_one & ~same_normal))
    assert fell >= 1 and c[2] >= 1, (fell, c[2])
    t.close()
