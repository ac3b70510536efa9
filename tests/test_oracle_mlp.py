"""Pins of the CPU oracle of the correction network (oracle/mlp_oracle.py).

The network and its training data have no reference counterpart; these checks
fix them by mathematics: finite differences, the quadrupole limit of a cell's
field, symmetry of the features, and a cover property of the band walk.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import mlp_oracle as mo
import oracle_lib as ol


def _random_params(seed):
    rng = np.random.default_rng(seed)
    p = mo.glorot(seed).astype(np.float64)
    # non-zero biases so that every bias path is exercised
    for a, b in ((mo.OFF_B1, mo.OFF_W2), (mo.OFF_B2, mo.OFF_W3), (mo.OFF_B3, mo.N_PARAMS)):
        p[a:b] = rng.normal(0, 0.3, b - a)
    return p


def test_bf16_rounding():
    one = np.float32(1.0)
    ulp = np.float32(2.0 ** -7)
    cases = {
        1.0: 1.0,
        1.0 + 2.0 ** -8: 1.0,                       # tie -> even mantissa
        1.0 + 3 * 2.0 ** -8: 1.0 + 2 * 2.0 ** -7,  # tie -> even mantissa
        1.0 + 2.0 ** -8 + 2.0 ** -20: 1.0 + 2.0 ** -7,
        -1.5: -1.5,
        0.0: 0.0,
    }
    for x, want in cases.items():
        assert mo.bf16(np.float32(x))[()] == np.float32(want), x
    rng = np.random.default_rng(1)
    x = (rng.normal(size=20000) * np.exp(rng.uniform(-20, 20, 20000))).astype(np.float32)
    b = mo.bf16(x)
    assert np.all(b.view(np.uint32) & 0xFFFF == 0)
    # nearest of the two bracketing BF16 values
    lo = (x.view(np.uint32) & 0xFFFF0000).view(np.float32)
    hi = ((x.view(np.uint32) & 0xFFFF0000) + 0x10000).view(np.float32)
    best = np.minimum(np.abs(x.astype(np.float64) - lo), np.abs(x.astype(np.float64) - hi))
    assert np.all(np.abs(x.astype(np.float64) - b) == best)
    assert one + ulp == np.float32(1.0078125)


def _random_basis(rng, m):
    b = rng.normal(size=(m, 16))
    b[:, 15] = 0.0
    return b


def test_loss_grad_finite_differences():
    rng = np.random.default_rng(2)
    p = _random_params(3)
    m = 37                                    # not a multiple of the 16-row tile
    x = rng.normal(size=(m, mo.K_IN))
    bs = _random_basis(rng, m)
    t = rng.normal(size=(m, mo.K_OUT))
    loss, g = mo.loss_grad(p, x, bs, t, "f64")
    blocks = [(mo.OFF_W1, mo.OFF_B1), (mo.OFF_B1, mo.OFF_W2), (mo.OFF_W2, mo.OFF_B2), (mo.OFF_B2, mo.OFF_W3),
              (mo.OFF_W3, mo.OFF_B3), (mo.OFF_B3, mo.N_PARAMS)]
    h = 1e-6
    for a, b in blocks:
        idx = rng.choice(np.arange(a, b), size=min(24, b - a), replace=False)
        for i in idx:
            pp = p.copy(); pp[i] += h
            pm = p.copy(); pm[i] -= h
            fd = (mo.loss_grad(pp, x, bs, t)[0] - mo.loss_grad(pm, x, bs, t)[0]) / (2 * h)
            assert abs(fd - g[i]) <= 1e-6 * max(1.0, abs(g[i])), (i, fd, g[i])
    # unused output columns (9..15) receive no gradient
    W3g = g[mo.OFF_W3:mo.OFF_B3].reshape(mo.K_HID, mo.K_OUTP)
    assert np.all(W3g[:, mo.K_OUT:] == 0) and np.all(g[mo.OFF_B3 + mo.K_OUT:] == 0)
    assert loss > 0


def test_rounding_modes_agree():
    rng = np.random.default_rng(4)
    p = _random_params(5)
    x = rng.normal(size=(64, mo.K_IN))
    b = _random_basis(rng, 64)
    t = rng.normal(size=(64, mo.K_OUT))
    l64, g64 = mo.loss_grad(p, x, b, t, "f64")
    l32, g32 = mo.loss_grad(p, x, b, t, "f32")
    assert abs(l32 - l64) < 1e-5 * l64
    assert np.linalg.norm(g32 - g64) < 1e-5 * np.linalg.norm(g64)
    lb, gb = mo.loss_grad(p, x, b, t, "bf16")
    assert np.linalg.norm(gb - g64) < 5e-2 * np.linalg.norm(g64)
    assert np.dot(gb, g64) > 0.99 * np.linalg.norm(gb) * np.linalg.norm(g64)


def _cluster(rng, k=200, scale=1.0):
    pos = rng.normal(size=(k, 3)) * scale * np.array([1.0, 0.6, 0.3])
    mass = rng.uniform(0.5, 1.5, k)
    return pos, mass


def test_field_jacobian_finite_differences():
    rng = np.random.default_rng(6)
    pos, mass = _cluster(rng)
    p = np.array([4.0, -3.0, 2.5])
    a, J = mo.field(p, pos, mass)
    h = 1e-5
    for j in range(3):
        e = np.zeros(3); e[j] = h
        fd = (mo.field(p + e, pos, mass)[0] - mo.field(p - e, pos, mass)[0]) / (2 * h)
        assert np.allclose(fd, J[:, j], rtol=1e-6, atol=1e-9 * np.abs(J).max())
    assert np.allclose(J, J.T)
    # traceless away from the sources (Laplace)
    assert abs(np.trace(J)) < 1e-12 * np.abs(J).max()


def test_residual_tends_to_quadrupole():
    rng = np.random.default_rng(7)
    pos, mass = _cluster(rng)
    cm, Q = mo.second_moments(pos, mass)
    errs = []
    for D in (20.0, 40.0, 80.0):
        p = cm + D * np.array([0.48, -0.6, 0.64])
        R, _ = mo.cell_residual(p, pos, mass)
        quad = mo.quadrupole_field(p, cm, Q)
        errs.append(np.linalg.norm(R - quad) / np.linalg.norm(quad))
    # the remainder is the octupole: relative error ~ 1/D
    assert errs[0] < 0.1
    assert errs[1] < 0.6 * errs[0] and errs[2] < 0.6 * errs[1]


def test_remainder_has_octupole_order():
    """cell_remainder = field - monopole - quadrupole: shrinking the cell about
    its mass centre by 2 divides it by about 8 (the octupole is cubic in the
    size); the same check on a copy that drops, or doubles, the quadrupole
    term gives about 4, so neither mistake passes."""
    rng = np.random.default_rng(8)
    pos, mass = _cluster(rng)
    cm, Q = mo.second_moments(pos, mass)
    p = cm + 30.0 * np.array([0.48, -0.6, 0.64])

    def shrunk(k):
        return cm + (pos - cm) / k

    def ratio(fn):
        r1 = np.linalg.norm(fn(p, shrunk(1.0), mass)[0])
        r2 = np.linalg.norm(fn(p, shrunk(2.0), mass)[0])
        return r1 / r2

    assert 7.0 < ratio(mo.cell_remainder) < 9.0
    assert 3.5 < ratio(mo.cell_residual) < 4.5                  # quadrupole dropped
    doubled = lambda pp, pos_, m_: (mo.cell_residual(pp, pos_, m_)[0]
                                    - 2.0 * mo.quadrupole_field(pp, *mo.second_moments(pos_, m_)), None)
    assert 3.5 < ratio(doubled) < 4.5
    # the Jacobian of the remainder against finite differences of its field
    R, J = mo.cell_remainder(p, pos, mass)
    h = 1e-3
    for j in range(3):
        e = np.zeros(3); e[j] = h
        fd = (mo.cell_remainder(p + e, pos, mass)[0] - mo.cell_remainder(p - e, pos, mass)[0]) / (2 * h)
        assert np.allclose(fd, J[:, j], rtol=1e-4, atol=1e-6 * np.abs(J).max())
    # order-2 targets are the remainder over M rho^3 / d^2
    lo, hi = pos.min(0), pos.max(0)
    f, b, t = mo.target_vector(p, pos, mass, lo, hi, order=2)
    _, _, scale, d = mo.cell_features(p, cm, mass.sum(), lo, hi, Q)
    assert np.allclose(t[:3] * scale * f[0], R, rtol=1e-12)
    assert np.allclose(t[3:] * scale * f[0] / (mo.K_JAC * d), [J[0, 0], J[1, 1], J[2, 2], J[0, 1], J[0, 2], J[1, 2]],
                       rtol=1e-12)


def _rotation(rng):
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    return q * np.sign(np.linalg.det(q))


def test_features_invariant_basis_covariant():
    rng = np.random.default_rng(8)
    pos, mass = _cluster(rng)
    pos = pos + np.array([0.3, -0.2, 0.1])     # box centre away from the mass centre
    p = np.array([5.0, 1.0, -2.0])

    def feats(P, X):
        cm, Q = mo.second_moments(X, mass)
        return mo.cell_features(P, cm, mass.sum(), X.min(0), X.max(0), Q)

    f0, b0, s0, d0 = feats(p, pos)
    assert abs(np.linalg.norm(b0[:3]) - 1) < 1e-12
    cm = (pos * mass[:, None]).sum(0) / mass.sum()
    r0 = mo.cell_radius(pos.min(0), pos.max(0), cm)
    assert np.isclose(f0[0], r0 / np.linalg.norm(cm - p)) and np.isclose(f0[1], f0[0] ** 2)
    assert np.isclose(d0, np.linalg.norm(cm - p)) and np.isclose(s0, mass.sum() * r0 ** 2 / d0 ** 4)
    # a symmetry of the axis-aligned box (signed axis permutation): features
    # unchanged, basis vectors and q transformed
    P = np.array([[0, 1, 0], [0, 0, -1], [-1, 0, 0]], float)
    f1, b1, s1, d1 = feats(P @ p, pos @ P.T)
    assert np.allclose(f1, f0) and np.isclose(s1, s0) and np.isclose(d1, d0)
    for k in (0, 3, 6):
        assert np.allclose(b1[k:k + 3], P @ b0[k:k + 3])
    Qm = lambda v: np.array([[v[0], v[3], v[4]], [v[3], v[1], v[5]], [v[4], v[5], v[2]]])
    assert np.allclose(Qm(b1[9:15]), P @ Qm(b0[9:15]) @ P.T)
    # predictions of any network turn with the cell
    prm = _random_params(9)
    y0 = mo.predict(prm, f0[None], b0[None])[0]
    y1 = mo.predict(prm, f1[None], b1[None])[0]
    assert np.allclose(y1[:3], P @ y0[:3])
    assert np.allclose(Qm(y1[3:]), P @ Qm(y0[3:]) @ P.T)


def test_quadrupole_coefficients_reproduce_quadrupole_field():
    """The basis spans the quadrupole term: its closed-form coefficients give
    the textbook quadrupole field and its finite-difference Jacobian."""
    rng = np.random.default_rng(11)
    pos, mass = _cluster(rng)
    cm, Q = mo.second_moments(pos, mass)
    for p in (np.array([6.0, -2.0, 3.0]), np.array([-1.0, 9.0, 0.5])):
        f, b, scale, d = mo.cell_features(p, cm, mass.sum(), pos.min(0), pos.max(0), Q)
        c = mo.quadrupole_coefficients(f)
        y = np.einsum("c,ck->k", c, mo.basis_matrix(b[None])[0])
        assert np.allclose(y[:3] * scale, mo.quadrupole_field(p, cm, Q), rtol=1e-12)
        J = np.zeros((3, 3))
        h = 1e-6 * d
        for j in range(3):
            e = np.zeros(3); e[j] = h
            J[:, j] = (mo.quadrupole_field(p + e, cm, Q) - mo.quadrupole_field(p - e, cm, Q)) / (2 * h)
        Jv = np.array([J[0, 0], J[1, 1], J[2, 2], J[0, 1], J[0, 2], J[1, 2]])
        assert np.allclose(y[3:] * scale / (mo.K_JAC * d), Jv, rtol=1e-6, atol=1e-9 * np.abs(Jv).max())
        # and the exact residual is close to it for this distant cell
        _, _, t = mo.target_vector(p, pos, mass, pos.min(0), pos.max(0))
        assert np.linalg.norm(t - y) < 0.3 * np.linalg.norm(y)


def _plummer_like(n, seed):
    rng = np.random.default_rng(seed)
    r = 1.0 / np.sqrt(rng.uniform(0.05, 1.0, n) ** (-2.0 / 3.0) - 1.0 + 1e-9)
    v = rng.normal(size=(n, 3))
    pos = v / np.linalg.norm(v, axis=1)[:, None] * r[:, None]
    mass = rng.uniform(0.5, 1.5, n) / n
    y = np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], np.zeros(3 * n)])
    return y, mass


@pytest.mark.parametrize("ratio", [1.25, 2.0])
def test_group_band_covers_every_body_once(ratio):
    n = 2048
    y, mass = _plummer_like(n, 9)
    h = ol.Heap(n)
    h.build(y, mass, ratio)
    a = h.arrays()
    for group in (0, 7, n // 64 - 1):
        far, taken, near = mo.group_band(n, group, a["cm"], a["crit_r2"], a["bmin"], a["bmax"], 0.2, 2.0)
        # the leaf slices of the taken nodes partition [n, 2n)
        cover = np.zeros(2 * n, int)
        for node in taken:
            depth = int(node).bit_length() - 1
            span = int(np.log2(n)) - depth
            cover[node << span:(node + 1) << span] += 1
        assert np.all(cover[n:] == 1)
        assert set(far) <= set(taken) and set(near) <= set(taken) and not set(far) & set(near)
        assert all(b < n for b in far + near)
        # near_k = 0: no near cells, the far set only grows
        far0, taken0, near0 = mo.group_band(n, group, a["cm"], a["crit_r2"], a["bmin"], a["bmax"], 0.2, 0.0)
        assert taken0 == taken and not near0 and set(far) <= set(far0)


def test_quadrupole_sum_improves_monopole():
    """Monopole + quadrupole of the taken cells is closer to the direct sum than
    the monopole alone, for every body of a group (small tree, large ratio)."""
    n = 1024
    y, mass = _plummer_like(n, 10)
    h = ol.Heap(n)
    h.build(y, mass, 3.0)
    a = h.arrays()
    pos, ms = a["cm"][n:], a["mass"][n:]
    far, taken, near = mo.group_band(n, 3, a["cm"], a["crit_r2"], a["bmin"], a["bmax"], 0.0, 1e9)
    assert not far and near
    worse = 0
    for leaf in range(n + 3 * 64, n + 4 * 64):
        x = a["cm"][leaf]
        exact = mo.field(x, pos, ms)[0]
        mono = sum(mo.field(x, a["cm"][c][None], np.array([a["mass"][c]]))[0] for c in taken)
        quad = mono.copy()
        for c in near:
            depth = int(c).bit_length() - 1
            s = int(np.log2(n)) - depth
            l0, l1 = (c << s) - n, ((c + 1) << s) - n
            _, Q = mo.second_moments(pos[l0:l1], ms[l0:l1])
            quad += mo.quadrupole_field(x, a["cm"][c], Q)
        worse += np.linalg.norm(quad - exact) > np.linalg.norm(mono - exact)
    assert worse <= 3


def _point_fields(x, R, M):
    """Exact monopole fields of point masses M[c] at R[c] at the point x."""
    return sum(mo.field(x, r[None, :], np.array([mm]))[0] for r, mm in zip(R, M))


def test_local_expansion_field_jacobian_and_error_order():
    """Expansions of point-mass fields about a centre: the field at the centre
    is the exact field, the Jacobian is the finite-difference one, and the error
    falls as |delta|^3 at second order and |delta|^4 at third -- halving the
    offsets divides the largest error by about 8 and 16, where dropping the
    second-order term (planted here in a copy of the coefficients) leaves a
    ratio near 4, and a sign error in the third-order term one near 8."""
    rng = np.random.default_rng(11)
    R = rng.normal(size=(6, 3)) * 2.0 + np.array([5.0, -4.0, 3.0])
    M = rng.uniform(0.1, 1.0, 6)
    f, J, D2, D3 = mo.local_expansion_tensors(R, M, 3)
    assert np.allclose(f, _point_fields(np.zeros(3), R, M), rtol=1e-13, atol=0)
    h = 1e-5
    Jfd = np.stack([(_point_fields(h * e, R, M) - _point_fields(-h * e, R, M)) / (2 * h) for e in np.eye(3)], axis=1)
    assert np.allclose(J, Jfd, rtol=1e-7, atol=1e-9 * np.abs(J).max())
    U = rng.normal(size=(12, 3))
    U /= np.linalg.norm(U, axis=1)[:, None]

    def worst(D2c, s, D3c=None):
        return max(np.linalg.norm(mo.local_expansion_eval(f, J, D2c, s * u[None, :], D3c)[0]
                                  - _point_fields(s * u, R, M)) for u in U)

    ratio = worst(D2, 0.2) / worst(D2, 0.1)
    assert 6.5 < ratio < 9.5, ratio
    planted = worst(np.zeros_like(D2), 0.2) / worst(np.zeros_like(D2), 0.1)
    assert 3.0 < planted < 5.0, planted
    ratio3 = worst(D2, 0.2, D3) / worst(D2, 0.1, D3)
    assert 13.0 < ratio3 < 19.0, ratio3
    assert worst(D2, 0.1, D3) < 0.2 * worst(D2, 0.1)
    planted3 = worst(D2, 0.2, -D3) / worst(D2, 0.1, -D3)
    assert 6.5 < planted3 < 9.5, planted3


def test_quadrupole_fields_matches_single_cell_form():
    rng = np.random.default_rng(12)
    P = rng.normal(size=(5, 3))
    cm = rng.normal(size=(4, 3)) + 6.0
    Q = rng.normal(size=(4, 6)) * 0.1
    want = np.array([sum(mo.quadrupole_field(p, c, q) for c, q in zip(cm, Q)) for p in P])
    assert np.allclose(mo.quadrupole_fields(P, cm, Q), want, rtol=1e-12, atol=1e-15)


def test_two_level_cell_sets_cover_every_body_once():
    """The two-level local walk's cell sets (super-group walk, then each of its
    16 groups from the list) cover every body of the tree exactly once per group,
    and the super-group walk does share far cells at this size."""
    n, ratio, near_k = 1 << 17, 1.0 / 0.6, 3.0
    y, m = _plummer_like(n, 40)
    h = ol.Heap(n)
    h.build(y, m, ratio)
    o = h.arrays()
    shared = 0
    for s in (3, 100):
        far, listed = mo.super_band(n, s, o["cm"], o["crit_r2"], o["bmin"], o["bmax"], near_k, ratio, shift=4)
        shared += len(far)
        for gsub in (0, 5):
            _, taken, _ = mo.group_band_from(n, 16 * s + gsub, listed, o["cm"], o["crit_r2"], o["bmin"], o["bmax"],
                                             near_k)
            cover = np.zeros(n, int)
            for node in far + taken:
                depth = int(node).bit_length() - 1
                span = int(np.log2(n)) - depth
                cover[(node << span) - n:((node + 1) << span) - n] += 1
            assert cover.min() == 1 and cover.max() == 1
    assert shared > 0
