"""The oracle's Morton radix tree (oracle_radix_build) against its definition.

The tree is pinned by quantities computed here independently of Karras'
parallel construction: the Morton codes from Python integers, the sort order,
the recursive definition of the binary radix tree (a node over the sorted keys
[a, b] splits after the last key whose bit at the highest position where keys a
and b differ is 0, each key extended by its position in the sorted order, as
Karras does for equal codes), and per node the mass,
box, mass centre and second moments summed over its bodies."""
import numpy as np
import pytest

import oracle_lib


def _morton(y, n):
    pos = np.stack([y[:n], y[n:2 * n], y[2 * n:3 * n]], 1)
    lo, hi = pos.min(0), pos.max(0)
    side = float((hi - lo).max())
    scale = 2097152.0 / side if side > 0 else 0.0
    cells = np.clip(np.floor((pos - lo) * scale), 0, 2097151).astype(np.int64)

    def spread(v):
        out = 0
        for b in range(21):
            out |= ((v >> b) & 1) << (3 * b)
        return out

    return [(spread(int(c[0])) << 2) | (spread(int(c[1])) << 1) | spread(int(c[2])) for c in cells]


def _definition(keys):
    """{(a, b): split} of every internal node of the radix tree of the sorted
    augmented keys (code << 32 | sorted position)."""
    out = {}
    stack = [(0, len(keys) - 1)]
    while stack:
        a, b = stack.pop()
        if a == b:
            continue
        bit = (keys[a] ^ keys[b]).bit_length() - 1
        s = a
        while s + 1 <= b and not (keys[s + 1] >> bit) & 1:
            s += 1
        out[(a, b)] = s
        stack += [(a, s), (s + 1, b)]
    return out


def _ranges(t, n):
    """(first, last) sorted leaf of every id, from the oracle's children."""
    first = np.zeros(2 * n, np.int64)
    last = np.zeros(2 * n, np.int64)
    first[n:] = np.arange(n)
    last[n:] = np.arange(n)
    order = []
    stack = [1]
    while stack:
        k = stack.pop()
        order.append(k)
        for c in t["child"][k]:
            if c < n:
                stack.append(int(c))
    for k in reversed(order):
        l, r = t["child"][k]
        first[k], last[k] = first[l], last[r]
    return first, last, order


def _state(kind, n, seed):
    rng = np.random.default_rng(seed)
    if kind == "normal":
        pos = rng.standard_normal((n, 3))
    elif kind == "clustered":
        pos = np.concatenate([rng.standard_normal((n // 2, 3)) * 1e-3, rng.standard_normal((n - n // 2, 3)) * 10])
    else:   # duplicates: every body on one of 5 points
        pos = rng.standard_normal((5, 3))[rng.integers(0, 5, n)]
    y = np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], np.zeros(3 * n)])
    return y, rng.random(n) + 0.1


@pytest.mark.parametrize("kind,n", [("normal", 2), ("normal", 3), ("normal", 257), ("clustered", 500),
                                    ("duplicates", 300)])
def test_radix_structure_matches_definition(kind, n):
    y, m = _state(kind, n, n)
    t = oracle_lib.radix_build(y, m, 1.5)
    codes = _morton(y, n)
    order = sorted(range(n), key=lambda i: (codes[i], i))
    assert list(t["order"]) == order
    keys = [(codes[b] << 32) | j for j, b in enumerate(order)]
    want = _definition(keys)
    first, last, visited = _ranges(t, n)
    assert sorted(visited) == list(range(1, n))          # n - 1 internal nodes, each reached once
    got = {}
    for k in visited:
        l, r = t["child"][k]
        assert t["parent"][l] == k and t["parent"][r] == k
        assert last[l] + 1 == first[r]
        got[(int(first[k]), int(last[k]))] = int(last[l])
    assert got == want
    assert (first[1], last[1]) == (0, n - 1)


@pytest.mark.parametrize("kind,n", [("normal", 400), ("clustered", 300), ("duplicates", 200)])
def test_radix_moments_by_brute_force(kind, n):
    y, m = _state(kind, n, 7 * n)
    ratio = 1.6
    t = oracle_lib.radix_build(y, m, ratio)
    pos = np.stack([y[:n], y[n:2 * n], y[2 * n:3 * n]], 1)[t["order"]]
    mass = m[t["order"]]
    first, last, visited = _ranges(t, n)
    for k in visited:
        sl = slice(first[k], last[k] + 1)
        p, w = pos[sl], mass[sl]
        M = w.sum()
        cm = (p * w[:, None]).sum(0) / M
        d = p - cm
        q = np.array([np.sum(w * d[:, 0] ** 2), np.sum(w * d[:, 1] ** 2), np.sum(w * d[:, 2] ** 2),
                      np.sum(w * d[:, 0] * d[:, 1]), np.sum(w * d[:, 0] * d[:, 2]), np.sum(w * d[:, 1] * d[:, 2])])
        assert t["mass"][k] == pytest.approx(M, rel=1e-13)
        assert np.allclose(t["cm"][k], cm, rtol=1e-12, atol=1e-12 * np.abs(p).max())
        assert np.array_equal(t["bmin"][k], p.min(0)) and np.array_equal(t["bmax"][k], p.max(0))
        # atol: rounding of the mass centres in the parallel-axis sums (nodes of
        # coincident bodies have moments of that size instead of 0)
        atol = M * (1e-9 * np.abs(d).max() ** 2 + 1e-28 * np.abs(p).max() ** 2) + 1e-300
        assert np.allclose(t["quad"][k], q, rtol=1e-9, atol=atol)
        lo, hi = p.min(0), p.max(0)
        rad = 0.5 * np.linalg.norm(hi - lo) + np.linalg.norm(0.5 * (hi + lo) - cm)
        # abs: the rounding of the mass centre inside a box of coincident bodies
        assert t["crit"][k] == pytest.approx((rad * ratio) ** 2, rel=1e-12, abs=(ratio * 1e-14 * np.abs(p).max()) ** 2)
    # leaves: the bodies themselves
    assert np.array_equal(t["cm"][n:], pos) and np.array_equal(t["mass"][n:], mass)
    assert np.all(t["crit"][n:] == 0.0) and np.all(t["quad"][n:] == 0.0)


@pytest.mark.parametrize("kind,n", [("normal", 300), ("clustered", 500), ("duplicates", 400)])
def test_radix_subtree_ids_within_range(kind, n):
    """The internal ids of the subtree of any node over the sorted leaves [a, b]
    lie in [a + 1, b + 1] (each internal node's index is one end of its key
    range), the property by which the GPU build holds a walk group's nodes in
    LDS slots numbered from the group's first leaf."""
    y, m = _state(kind, n, 3 * n)
    t = oracle_lib.radix_build(y, m, 1.5)
    first, last, visited = _ranges(t, n)
    for k in visited:
        assert k - 1 in (first[k], last[k])           # index k - 1 is one end of the range
        stack = [k]
        while stack:
            j = stack.pop()
            assert first[k] + 1 <= j <= last[k] + 1
            stack += [int(c) for c in t["child"][j] if c < n]


@pytest.mark.parametrize("kind,n", [("normal", 257), ("clustered", 500), ("duplicates", 300)])
def test_radix_children_from_split(kind, n):
    """The children of every internal node follow from its split s alone: leaf
    n + s or internal id s + 1 on the left, leaf n + s + 1 or internal id s + 2 on
    the right (the walks store a node's children as s + 1 and two leaf flags)."""
    y, m = _state(kind, n, 5 * n)
    t = oracle_lib.radix_build(y, m, 1.5)
    for k in range(1, n):
        l, r = (int(c) for c in t["child"][k])
        g = l if l < n else l - n + 1                   # split + 1
        assert l == (n + g - 1 if l >= n else g)
        assert r == (n + g if r >= n else g + 1)
        # decode of the stack entry the walks push
        e = g | ((l >= n) << 30) | ((r >= n) << 31)
        gd = e & 0x3FFFFFFF
        assert ((n + gd - 1) if (e >> 30) & 1 else gd) == l
        assert ((n + gd) if e >> 31 else gd + 1) == r
