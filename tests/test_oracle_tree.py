"""Pins the oracle's heap kd-tree (nbody/nbody_space_heap*.cpp) and heap index
algebra (nbody/nbody_space_heap_func_priv.h) with the reference's KATs, an
independent statement of the median split by full sorting, structural
invariants and the direct-sum limit of the traversal."""
import numpy as np
import pytest

import oracle_lib
from nbody_amd import nbody_data

L = oracle_lib.lib


def test_heap_func_kats():
    """test/engine/test_nbody_engine.cpp:1107-1147 (tree of 16 nodes)."""
    lib = L()
    assert lib.oracle_heap_next_down(8) == 9
    assert lib.oracle_heap_next_down(9) == 5
    for idx, exp in [(8, 9), (9, 5), (5, 3), (11, 3), (3, 1), (7, 1), (15, 1)]:
        assert lib.oracle_heap_skip_idx(idx) == exp
    for idx, exp in [(1, 2), (8, 9), (9, 5), (15, 1)]:
        assert lib.oracle_heap_next_up(idx, 16) == exp
    for idx in range(1, 16):
        assert lib.oracle_heap_parent(lib.oracle_heap_left(idx)) == idx
        assert lib.oracle_heap_parent(lib.oracle_heap_right(idx)) == idx
        assert lib.oracle_heap_left(idx) + 1 == lib.oracle_heap_right(idx)
        # next_down via ffs, as the CUDA code computes it (test_nbody_engine.cpp:1085-1102)
        inv = ~idx & 0xFFFFFFFFFFFFFFFF
        ffs = (inv & -inv).bit_length()
        assert lib.oracle_heap_next_down(idx) == (idx >> (ffs - 1)) + 1


def _preorder(size):
    order = []

    def rec(i):
        if i >= size:
            return
        order.append(i)
        rec(2 * i)
        rec(2 * i + 1)
    rec(1)
    return order


@pytest.mark.parametrize("k", range(1, 12))
def test_heap_func_vs_pointer_preorder(k):
    """next_up = preorder successor, skip_idx = successor after the subtree, for every node."""
    lib = L()
    size = 2 << k
    pre = _preorder(size)
    pos = {v: i for i, v in enumerate(pre)}
    for i in pre:
        nxt = pre[pos[i] + 1] if pos[i] + 1 < len(pre) else 1
        assert lib.oracle_heap_next_up(i, size) == nxt
        # end of subtree of i in preorder
        last = i
        while 2 * last + 1 < size:
            last = 2 * last + 1
        after = pre[pos[last] + 1] if pos[last] + 1 < len(pre) else 1
        assert lib.oracle_heap_skip_idx(i) == after


def _sorted_definition(pos, n):
    """The median split stated by sorting: left child = count/2 smallest along
    axis depth % 3 (ties by index).  Returns body per leaf, leaves n..2n-1."""
    leaf = np.empty(n, dtype=np.int64)

    def rec(ind, idx, dim):
        if len(ind) == 1:
            leaf[idx - n] = ind[0]
            return
        order = ind[np.lexsort((ind, pos[ind, dim]))]
        h = len(ind) // 2
        rec(order[:h], 2 * idx, (dim + 1) % 3)
        rec(order[h:], 2 * idx + 1, (dim + 1) % 3)
    rec(np.arange(n), 1, 0)
    return leaf


@pytest.mark.parametrize("kind,n", [("ball", 2), ("plummer", 64), ("cube", 1024), ("two_plummer", 4096)])
def test_build_matches_sorting_definition(kind, n):
    d = nbody_data()
    d.set_quiet(True)
    d.generate(kind, n, 13)
    y, m = oracle_lib.state_from_data(d)
    h = oracle_lib.Heap(n)
    h.build(y, m, 2.0)
    arr = h.arrays()
    pos = y[:3 * n].reshape(3, n).T.copy()
    assert np.array_equal(arr["body_n"][n:], _sorted_definition(pos, n))
    assert np.all(arr["body_n"][1:n] == -1)


def test_build_invariants():
    n = 2048
    d = nbody_data()
    d.set_quiet(True)
    d.generate("plummer", n, 4)
    y, m = oracle_lib.state_from_data(d)
    pos = y[:3 * n].reshape(3, n).T
    h = oracle_lib.Heap(n)
    h.build(y, m, 3.0)
    a = h.arrays()
    body = a["body_n"]
    assert sorted(body[n:]) == list(range(n))
    # leaves: moments are the body itself, crit 0
    assert np.array_equal(a["cm"][n:], pos[body[n:]])
    assert np.array_equal(a["mass"][n:], m[body[n:]])
    assert np.all(a["crit_r2"][n:] == 0)
    for idx in range(1, n):
        depth = idx.bit_length() - 1
        lo, hi = idx << (np.log2(n).astype(int) - depth), (idx + 1) << (np.log2(n).astype(int) - depth)
        bodies = body[lo:hi]
        # subtree = contiguous leaf slice; box = exact min/max; mass ~ sum
        assert np.array_equal(a["bmin"][idx], pos[bodies].min(axis=0))
        assert np.array_equal(a["bmax"][idx], pos[bodies].max(axis=0))
        assert a["mass"][idx] == pytest.approx(m[bodies].sum(), rel=1e-14)
        cm = (pos[bodies] * m[bodies, None]).sum(axis=0) / m[bodies].sum()
        assert np.allclose(a["cm"][idx], cm, rtol=1e-12, atol=1e-14)
        half = len(bodies) // 2
        ax = depth % 3
        assert pos[bodies[:half], ax].max() <= pos[bodies[half:], ax].min()
        box_r = 0.5 * np.linalg.norm(a["bmax"][idx] - a["bmin"][idx]) + \
            np.linalg.norm((a["bmax"][idx] + a["bmin"][idx]) / 2 - a["cm"][idx])
        assert a["crit_r2"][idx] == pytest.approx((box_r * 3.0) ** 2, rel=1e-14)


def test_two_body_node_closed_form():
    y = np.array([0.0, 4.0, 0.0, 0.0, 0.0, 0.0, 0, 0, 0, 0, 0, 0])
    m = np.array([1.0, 3.0])
    h = oracle_lib.Heap(2)
    h.build(y, m, 2.0)
    a = h.arrays()
    assert list(a["body_n"][2:]) == [0, 1]
    assert a["mass"][1] == 4.0
    assert np.array_equal(a["cm"][1], [3.0, 0.0, 0.0])
    # r = |box|/2 + |box centre - cm| = 2 + 1, crit = (2 r)^2
    assert a["crit_r2"][1] == 36.0


def test_traversal_ratio_1e8_is_direct_sum_in_leaf_order():
    """Every internal node opens, every other leaf is accepted: the sum is the
    direct sum taken in heap-leaf order, bit for bit."""
    n = 256
    d = nbody_data()
    d.set_quiet(True)
    d.generate("ball", n, 5)
    y, m = oracle_lib.state_from_data(d)
    pos = y[:3 * n].reshape(3, n).T
    h = oracle_lib.Heap(n)
    h.build(y, m, 1e8)
    body = h.arrays()["body_n"][n:]
    for tgt in (0, 17, n - 1):
        v1 = pos[tgt]
        fx = fy = fz = 0.0
        for b in body:
            if b == tgt:
                continue
            dx, dy, dz = v1 - pos[b]
            r2 = dx * dx + dy * dy + dz * dz
            r2 = max(r2, 1e-8)
            c = ((-1.0 * m[tgt]) * m[b]) / (r2 * np.sqrt(r2))
            fx += dx * c
            fy += dy * c
            fz += dz * c
        for stackless in (False, True):
            assert np.array_equal(h.traverse(v1, m[tgt], stackless), [fx, fy, fz])


@pytest.mark.parametrize("ratio", [0.5, 1.25, 2.0, 3.1623, 10.0])
def test_heap_equals_stackless(ratio):
    n = 1024
    d = nbody_data()
    d.set_quiet(True)
    d.generate("plummer", n, 6)
    y, m = oracle_lib.state_from_data(d)
    h = oracle_lib.Heap(n)
    h.build(y, m, ratio)
    assert np.array_equal(h.fcompute(y, m, 1, 1), h.fcompute(y, m, 2, 1))
    assert np.array_equal(h.fcompute(y, m, 2, 0), h.fcompute(y, m, 2, 1))    # cycle == nested_tree


def test_far_target_sees_monopole():
    """A target far outside the root's opening radius interacts with the root only."""
    n = 64
    d = nbody_data()
    d.set_quiet(True)
    d.generate("ball", n, 7)
    y, m = oracle_lib.state_from_data(d)
    h = oracle_lib.Heap(n)
    h.build(y, m, 0.5)
    a = h.arrays()
    v1 = np.array([1e3, 0.0, 0.0])
    f = h.traverse(v1, 2.0)
    out = np.empty(3)
    L().oracle_force(oracle_lib.p(v1), oracle_lib.p(np.ascontiguousarray(a["cm"][1])), 2.0, a["mass"][1],
                     oracle_lib.p(out))
    assert np.array_equal(f, out)


def test_bh_error_decreases_with_ratio():
    n = 2048
    d = nbody_data()
    d.set_quiet(True)
    d.generate("plummer", n, 8)
    y, m = oracle_lib.state_from_data(d)
    a0 = oracle_lib.fcompute_simple(y, m, 0)[3 * n:].reshape(3, n)
    errs = []
    for ratio in (1.0, 2.0, 4.0, 10.0):
        h = oracle_lib.Heap(n)
        h.build(y, m, ratio)
        a = h.fcompute(y, m)[3 * n:].reshape(3, n)
        errs.append(np.sqrt(np.mean((np.linalg.norm(a - a0, axis=0) / np.linalg.norm(a0, axis=0)) ** 2)))
    # monopole cells: error falls roughly as lambda^-2
    assert errs[0] > errs[1] > errs[2] > errs[3]
    assert errs[3] < 2e-3
    assert errs[1] / errs[2] > 2.5 and errs[2] / errs[3] > 5.0


def test_rebuild_equals_build_when_order_is_kept():
    n = 512
    d = nbody_data()
    d.set_quiet(True)
    d.generate("plummer", n, 9)
    y, m = oracle_lib.state_from_data(d)
    h1 = oracle_lib.Heap(n)
    h1.build(y, m, 2.0)
    y2 = y.copy()
    y2[:3 * n] *= 1.5                               # positive scaling keeps every ordering
    h1.rebuild(y2, 2.0)
    h2 = oracle_lib.Heap(n)
    h2.build(y2, m, 2.0)
    a1, a2 = h1.arrays(), h2.arrays()
    for k in ("body_n", "mass", "crit_r2", "cm", "bmin", "bmax"):
        assert np.array_equal(a1[k][1:], a2[k][1:]), k


def test_task_build_and_sampled_walks_equal_the_serial_engine():
    """The OpenMP-task build (timed as the CPU baseline) gives the serial
    build's heap bit for bit, and the sampled walks give the fcompute rows of
    the same bodies (both layouts)."""
    n = 1 << 13
    d = nbody_data()
    d.set_quiet(True)
    d.generate("plummer", n, 9)
    y, m = oracle_lib.state_from_data(d)
    h1, h2 = oracle_lib.Heap(n), oracle_lib.Heap(n)
    h1.build(y, m, 2.0)
    h2.build_omp(y, m, 2.0, threads=4)
    a1, a2 = h1.arrays(), h2.arrays()
    for k in a1:
        assert np.array_equal(a1[k], a2[k]), k
    idx = np.random.default_rng(1).choice(n, 300, replace=False)
    for layout in (1, 2):
        f = h1.fcompute(y, m, layout, 1)
        acc = h2.accel_targets(y, m, idx, layout, threads=3)
        want = f[3 * n:].reshape(3, n).T[idx]
        assert np.array_equal(acc, want)
