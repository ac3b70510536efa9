"""The learned Barnes-Hut correction on the GPU against the CPU oracle
(oracle/mlp_oracle.py): BF16 matrix-core forward and hand-written backward,
training, the band-cell sampler, the corrected group walk, and the gain of the
trained network over the plain walk at the same opening ratio."""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from hip_capi import Ctx
from nbody_amd import native, nbody_data
from oracle import mlp_oracle as mo

pytestmark = pytest.mark.gpu

F32 = np.float32


@pytest.fixture(scope="module")
def ctx():
    c = Ctx(0)
    yield c
    c.close()


class Net:
    def __init__(self, ctx, seed=1):
        self.ctx = ctx
        self.h = C.c_void_p()
        native.check(ctx.lib.nbh_mlp_create(ctx.h, seed, C.byref(self.h)), "mlp_create")

    def close(self):
        native.check(self.ctx.lib.nbh_mlp_destroy(self.h), "mlp_destroy")

    def set(self, p):
        p = np.ascontiguousarray(p, F32)
        native.check(self.ctx.lib.nbh_mlp_set_params(self.ctx.h, self.h, p.ctypes.data_as(C.c_void_p)), "set")

    def get(self):
        p = np.empty(mo.N_PARAMS, F32)
        native.check(self.ctx.lib.nbh_mlp_get_params(self.ctx.h, self.h, p.ctypes.data_as(C.c_void_p)), "get")
        return p

    def forward(self, x):
        m = x.shape[0]
        dx = self.ctx.put_raw(np.ascontiguousarray(x, F32))
        do = self.ctx.zeros_raw(F32, m * mo.K_OUT)
        native.check(self.ctx.lib.nbh_mlp_forward(self.ctx.h, self.h, dx.ptr, m, do.ptr), "forward")
        self.ctx.sync()
        return do.read_raw(F32, m * mo.K_OUT).reshape(m, mo.K_OUT)

    def gradient(self, dx, db, dt, m):
        g = np.empty(mo.N_PARAMS, F32)
        loss = C.c_float()
        native.check(self.ctx.lib.nbh_mlp_gradient(self.ctx.h, self.h, dx.ptr, db.ptr, dt.ptr, m,
                                                   g.ctypes.data_as(C.c_void_p), C.byref(loss)), "gradient")
        return g.astype(np.float64), loss.value

    def train(self, dx, db, dt, m, epochs, batch, lr):
        el = np.empty(epochs, F32)
        native.check(self.ctx.lib.nbh_mlp_train(self.ctx.h, self.h, dx.ptr, db.ptr, dt.ptr, m, epochs, batch, lr,
                                                el.ctypes.data_as(C.c_void_p)), "train")
        return el


def _random_basis(rng, m):
    b = rng.normal(size=(m, 16)).astype(F32)
    b[:, 15] = 0.0
    return b


def _random_params(seed, bias=0.3):
    rng = np.random.default_rng(seed)
    p = mo.glorot(seed)
    for a, b in ((mo.OFF_B1, mo.OFF_W2), (mo.OFF_B2, mo.OFF_W3), (mo.OFF_B3, mo.OFF_B3 + mo.K_OUT)):
        p[a:b] = rng.normal(0, bias, b - a)
    return p.astype(F32)


def test_param_count_and_roundtrip(ctx):
    cnt = C.c_int()
    native.check(ctx.lib.nbh_mlp_param_count(C.byref(cnt)))
    assert cnt.value == mo.N_PARAMS
    net = Net(ctx, 7)
    p0 = net.get()
    assert np.all(np.isfinite(p0)) and np.abs(p0[mo.OFF_W1:mo.OFF_B1]).max() > 0
    W3 = p0[mo.OFF_W3:mo.OFF_B3].reshape(mo.K_HID, mo.K_OUTP)
    assert np.all(W3[:, mo.K_OUT:] == 0)          # unused output columns start at zero
    p = _random_params(3)
    net.set(p)
    assert np.array_equal(net.get(), p)
    net.close()


@pytest.mark.parametrize("m", [1, 16, 1000])
def test_forward_matches_oracle(ctx, m):
    rng = np.random.default_rng(m)
    net = Net(ctx)
    p = _random_params(11)
    net.set(p)
    x = rng.normal(size=(m, mo.K_IN)).astype(F32)
    got = net.forward(x)
    want, _ = mo.forward(p, x, "bf16")
    exact, _ = mo.forward(p.astype(np.float64), x.astype(np.float64), "f64")
    scale = np.abs(want).max()
    # same BF16 operand rounding: differences come from FP32 summation order
    # and tanh, and rare BF16 rounding flips of hidden activations
    assert np.abs(got - want).max() <= 4e-3 * scale, np.abs(got - want).max() / scale
    assert np.abs(got - want).mean() <= 2e-4 * scale
    # BF16 operands vs float64: a few parts in a thousand
    assert np.linalg.norm(got - exact) <= 3e-2 * np.linalg.norm(exact)
    net.close()


@pytest.mark.parametrize("m", [16, 333, 4096])
def test_gradient_matches_oracle(ctx, m):
    """Gradient of the hand-written backward kernels against the oracle's
    (float64 and BF16-emulating) backpropagation, itself pinned by finite
    differences."""
    rng = np.random.default_rng(100 + m)
    net = Net(ctx)
    p = _random_params(12)
    net.set(p)
    x = rng.normal(size=(m, mo.K_IN)).astype(F32)
    bs = _random_basis(rng, m)
    t = rng.normal(size=(m, mo.K_OUT)).astype(F32)
    dx, db, dt = ctx.put_raw(x), ctx.put_raw(bs), ctx.put_raw(t)
    g, loss = net.gradient(dx, db, dt, m)
    lb, gb = mo.loss_grad(p, x, bs, t, "bf16")
    l64, g64 = mo.loss_grad(p.astype(np.float64), x.astype(np.float64), bs.astype(np.float64),
                            t.astype(np.float64), "f64")
    assert abs(loss - lb) <= 2e-3 * lb
    assert np.linalg.norm(g - gb) <= 1e-2 * np.linalg.norm(gb), np.linalg.norm(g - gb) / np.linalg.norm(gb)
    assert np.linalg.norm(g - g64) <= 5e-2 * np.linalg.norm(g64)
    assert np.dot(g, g64) >= 0.995 * np.linalg.norm(g) * np.linalg.norm(g64)
    # per block as well (a transposed or misplaced block would fail here)
    for a, b in ((mo.OFF_W1, mo.OFF_B1), (mo.OFF_B1, mo.OFF_W2), (mo.OFF_W2, mo.OFF_B2), (mo.OFF_B2, mo.OFF_W3),
                 (mo.OFF_W3, mo.OFF_B3), (mo.OFF_B3, mo.N_PARAMS)):
        assert np.linalg.norm(g[a:b] - gb[a:b]) <= 2e-2 * np.linalg.norm(gb[a:b]) + 1e-7, (a, b)
    net.close()


def test_training_reduces_loss(ctx):
    rng = np.random.default_rng(5)
    m = 16384
    x = rng.normal(size=(m, mo.K_IN)).astype(F32)
    bs = _random_basis(rng, m) * F32(0.5)
    A = rng.normal(size=(mo.K_IN, mo.K_OUT)) / 4
    c_true = np.tanh(x @ A) + 0.1 * np.sin(x[:, :mo.K_OUT] * 2)
    t = np.einsum("mc,mck->mk", c_true, mo.basis_matrix(bs.astype(np.float64))).astype(F32)
    dx, db, dt = ctx.put_raw(x), ctx.put_raw(bs), ctx.put_raw(t)
    net = Net(ctx, 3)
    _, l0 = net.gradient(dx, db, dt, m)
    el = net.train(dx, db, dt, m, 20, 512, 3e-3)
    _, l1 = net.gradient(dx, db, dt, m)
    assert np.all(np.isfinite(el))
    assert l1 < 0.1 * l0, (l0, l1, el)
    # the trained parameters mean the same thing to the oracle
    pout = mo.predict(net.get(), x[:2048], bs[:2048], "bf16")
    assert abs(np.sum((pout - t[:2048]) ** 2) / 2048 - l1) < 0.25 * l1 + 1e-4
    net.close()


# ---------------------------------------------------------------------------
# tree side
# ---------------------------------------------------------------------------
class Samples:
    def __init__(self, total, dx, db, dt, x, b, t, cells):
        self.total, self.dx, self.db, self.dt = total, dx, db, dt
        self.x, self.b, self.t, self.cells = x, b, t, cells


class Tree:
    def __init__(self, ctx, y, m, ratio):
        self.ctx, self.n = ctx, len(m)
        self.h = C.c_void_p()
        native.check(ctx.lib.nbh_tree_create(ctx.h, self.n, C.byref(self.h)), "tree_create")
        self.dy, self.dm = ctx.put(y), ctx.put(m)
        native.check(ctx.lib.nbh_tree_build(ctx.h, self.h, self.dy.ptr, self.dm.ptr, ratio), "build")

    def close(self):
        self.ctx.lib.nbh_tree_destroy(self.h)

    def arrays(self):
        s = 2 * self.n
        cm = np.empty(3 * s); mass = np.empty(s); crit = np.empty(s); bmin = np.empty(3 * s); bmax = np.empty(3 * s)
        P = lambda a: a.ctypes.data_as(C.c_void_p)
        native.check(self.ctx.lib.nbh_tree_get(self.ctx.h, self.h, None, P(cm), P(mass), P(crit), P(bmin), P(bmax)))
        return dict(cm=cm.reshape(-1, 3), mass=mass, crit=crit, bmin=bmin.reshape(-1, 3), bmax=bmax.reshape(-1, 3))

    def samples(self, groups, rho, near_k, capacity):
        g = np.ascontiguousarray(groups, np.uint32)
        dx = self.ctx.zeros_raw(F32, max(1, capacity) * mo.K_IN)
        db = self.ctx.zeros_raw(F32, max(1, capacity) * 16)
        dt = self.ctx.zeros_raw(F32, max(1, capacity) * mo.K_OUT)
        dc = self.ctx.zeros_raw(np.uint32, max(1, capacity) * 2)
        cnt = C.c_size_t()
        native.check(self.ctx.lib.nbh_bh_band_samples(self.ctx.h, self.h, g.ctypes.data_as(C.c_void_p), len(g),
                                                      rho, near_k, dx.ptr, db.ptr, dt.ptr, dc.ptr, capacity,
                                                      C.byref(cnt)), "samples")
        k = min(cnt.value, capacity)
        return Samples(cnt.value, dx, db, dt, dx.read_raw(F32, k * mo.K_IN).reshape(k, mo.K_IN),
                       db.read_raw(F32, k * 16).reshape(k, 16), dt.read_raw(F32, k * mo.K_OUT).reshape(k, mo.K_OUT),
                       dc.read_raw(np.uint32, 2 * k).reshape(k, 2))

    def super_samples(self, near_k, supers, capacity):
        sel = np.ascontiguousarray(supers, np.uint32)
        dx = self.ctx.zeros_raw(F32, max(1, capacity) * mo.K_IN)
        db = self.ctx.zeros_raw(F32, max(1, capacity) * 16)
        dt = self.ctx.zeros_raw(F32, max(1, capacity) * mo.K_OUT)
        dc = self.ctx.zeros_raw(np.uint32, max(1, capacity) * 2)
        cnt = C.c_size_t()
        native.check(self.ctx.lib.nbh_bh_super_samples(self.ctx.h, self.h, near_k, sel.ctypes.data_as(C.c_void_p),
                                                       len(sel), dx.ptr, db.ptr, dt.ptr, dc.ptr, capacity,
                                                       C.byref(cnt)), "super samples")
        k = min(cnt.value, capacity)
        return Samples(cnt.value, dx, db, dt, dx.read_raw(F32, k * mo.K_IN).reshape(k, mo.K_IN),
                       db.read_raw(F32, k * 16).reshape(k, 16), dt.read_raw(F32, k * mo.K_OUT).reshape(k, mo.K_OUT),
                       dc.read_raw(np.uint32, 2 * k).reshape(k, 2))

    def local2(self, near_k, net=None):
        df = self.ctx.alloc(6 * self.n)
        if net is None:
            native.check(self.ctx.lib.nbh_bh_fcompute_group_local2(self.ctx.h, self.h, self.dy.ptr, df.ptr, near_k),
                         "local2")
        else:
            native.check(self.ctx.lib.nbh_bh_fcompute_group_local2_mlp(self.ctx.h, self.h, net.h, self.dy.ptr,
                                                                        df.ptr, near_k), "local2_mlp")
        return df.read()

    def group(self):
        df = self.ctx.alloc(6 * self.n)
        native.check(self.ctx.lib.nbh_bh_fcompute_group(self.ctx.h, self.h, self.dy.ptr, df.ptr), "group")
        return df.read()

    def group_mlp(self, net, rho, near_k):
        df = self.ctx.alloc(6 * self.n)
        native.check(self.ctx.lib.nbh_bh_fcompute_group_mlp(self.ctx.h, self.h, net.h, self.dy.ptr, df.ptr, rho,
                                                            near_k), "group_mlp")
        return df.read()

    def counters(self):
        c = (C.c_uint64 * 5)()
        native.check(self.ctx.lib.nbh_bh_walk_counters(self.ctx.h, self.h, c))
        return list(c)


def _state(kind, n, seed):
    d = nbody_data()
    d.set_quiet(True)
    d.generate(kind, n, seed)
    return oracle_lib.state_from_data(d)


def _leaf_slice(n, node):
    depth = int(node).bit_length() - 1
    span = int(np.log2(n)) - depth
    return node << span, (node + 1) << span


def test_band_samples_match_oracle(ctx):
    n, ratio, rho, near_k = 4096, 1.25, 0.2, 2.0
    y, m = _state("plummer", n, 21)
    t = Tree(ctx, y, m, ratio)
    a = t.arrays()
    groups = np.arange(n // 64)
    smp = t.samples(groups, rho, near_k, 1 << 20)
    total, x, bs, tg, cells = smp.total, smp.x, smp.b, smp.t, smp.cells
    assert total == len(cells) > 0
    want = set()
    for g in groups:
        far, _, _ = mo.group_band(n, int(g), a["cm"], a["crit"], a["bmin"], a["bmax"], rho, near_k)
        want |= {(int(g), int(c)) for c in far}
    got = {(int(g), int(c)) for g, c in cells}
    assert got == want and len(got) == len(cells)
    for k in range(len(cells)):
        g, c = int(cells[k, 0]), int(cells[k, 1])
        gn = n // 64 + g
        p = 0.5 * (a["bmin"][gn] + a["bmax"][gn])
        l0, l1 = _leaf_slice(n, c)
        fx, fb, ft = mo.target_vector(p, a["cm"][l0:l1], a["mass"][l0:l1], a["bmin"][c], a["bmax"][c])
        assert np.allclose(x[k], fx, rtol=1e-5, atol=1e-6 * np.abs(fx).max()), (k, x[k], fx)
        assert np.allclose(bs[k], fb, rtol=1e-5, atol=1e-6 * np.abs(fb).max()), (k, bs[k], fb)
        assert np.allclose(tg[k], ft, rtol=1e-4, atol=1e-4 * np.abs(ft).max()), (k, tg[k], ft)
    # a smaller capacity: the same count, only `capacity` rows written
    smp2 = t.samples(groups[:8], rho, near_k, 5)
    total2, cells2 = smp2.total, smp2.cells
    assert len(cells2) == 5 and total2 == sum(1 for g, _ in want if g < 8)
    assert {(int(g), int(c)) for g, c in cells2} <= want
    t.close()


def test_band_samples_reject_bad_group(ctx):
    n = 1024
    y, m = _state("ball", n, 2)
    t = Tree(ctx, y, m, 1.25)
    g = np.array([0, n // 64], np.uint32)
    cnt = C.c_size_t()
    st = ctx.lib.nbh_bh_band_samples(ctx.h, t.h, g.ctypes.data_as(C.c_void_p), 2, 0.2, 2.0, None, None, None,
                                     None, 0, C.byref(cnt))
    assert st != 0 and b"out of range" in ctx.lib.nbh_last_error()
    t.close()


def test_zero_network_without_near_cells_is_group_walk(ctx):
    n = 1 << 14
    y, m = _state("plummer", n, 22)
    t = Tree(ctx, y, m, 1.25)
    plain = t.group()
    c_plain = t.counters()
    net = Net(ctx)
    net.set(np.zeros(mo.N_PARAMS, F32))
    got = t.group_mlp(net, 0.2, 0.0)
    c = t.counters()
    assert np.array_equal(got, plain)
    assert c[:3] == c_plain[:3] and c[3] > 0 and c[4] == 0
    net.close()
    t.close()


def _body_order(ctx, t):
    body = np.empty(2 * t.n, np.int32)
    native.check(ctx.lib.nbh_tree_get(ctx.h, t.h, body.ctypes.data_as(C.c_void_p), None, None, None, None, None))
    return body


def _oracle_correction(n, a, body, p, rho, near_k, use_net=True):
    """Per-body correction (original body order): quadrupole terms of the near
    cells plus the network's expansion of the far cells about the group centre."""
    want = np.zeros((n, 3))
    for g in range(n // 64):
        gn = n // 64 + g
        cen = 0.5 * (a["bmin"][gn] + a["bmax"][gn])
        far, _, near = mo.group_band(n, g, a["cm"], a["crit"], a["bmin"], a["bmax"], rho, near_k)
        S = np.zeros(mo.K_OUT)
        if far and use_net:
            feats, basis, w0, w1 = [], [], [], []
            for c in far:
                l0, l1 = _leaf_slice(n, c)
                _, Q = mo.second_moments(a["cm"][l0:l1], a["mass"][l0:l1])
                f, b, scale, d = mo.cell_features(cen, a["cm"][c], a["mass"][c], a["bmin"][c], a["bmax"][c], Q)
                feats.append(f); basis.append(b); w0.append(scale); w1.append(scale / (mo.K_JAC * d))
            out = mo.predict(p, np.array(feats, F32), np.array(basis, F32), "bf16").astype(np.float64)
            S[:3] = (out[:, :3] * np.array(w0)[:, None]).sum(0)
            S[3:] = (out[:, 3:] * np.array(w1)[:, None]).sum(0)
        J = np.array([[S[3], S[6], S[7]], [S[6], S[4], S[8]], [S[7], S[8], S[5]]])
        Qs = []
        for c in near:
            l0, l1 = _leaf_slice(n, c)
            Qs.append(mo.second_moments(a["cm"][l0:l1], a["mass"][l0:l1])[1])
        for leaf in range(n + 64 * g, n + 64 * g + 64):
            x = a["cm"][leaf]
            v = S[:3] + J @ (x - cen)
            for c, Q in zip(near, Qs):
                v = v + mo.quadrupole_field(x, a["cm"][c], Q)
            want[body[leaf]] = v
    return want


def test_near_quadrupole_matches_oracle(ctx):
    n, ratio, near_k = 4096, 1.25, 2.0
    y, m = _state("plummer", n, 24)
    t = Tree(ctx, y, m, ratio)
    a = t.arrays()
    net = Net(ctx)
    net.set(np.zeros(mo.N_PARAMS, F32))
    plain = t.group()
    got = t.group_mlp(net, 0.0, near_k)
    assert t.counters()[4] > 0
    delta = (got - plain)[3 * n:].reshape(3, n).T
    want = _oracle_correction(n, a, _body_order(ctx, t), None, 0.0, near_k, use_net=False)
    err = np.linalg.norm(delta - want, axis=1) / np.abs(want).max()
    assert err.max() < 1e-3, err.max()
    net.close()
    t.close()


def test_quadrupole_walk_equals_all_near_cells(ctx):
    """nbh_bh_fcompute_group_quad == the corrected walk with every taken cell
    near and a zero network (same cells, same order, same arithmetic), and it
    is more accurate than the monopole walk at the same ratio."""
    n, ratio = 1 << 14, 2.0
    y, m = _state("plummer", n, 25)
    t = Tree(ctx, y, m, ratio)
    df = ctx.alloc(6 * n)
    native.check(ctx.lib.nbh_bh_fcompute_group_quad(ctx.h, t.h, t.dy.ptr, df.ptr, float("inf")), "quad")
    quad = df.read()
    assert t.counters()[2] == 0                 # no stack spills: identical cell sets
    net = Net(ctx)
    net.set(np.zeros(mo.N_PARAMS, F32))
    ref = t.group_mlp(net, 0.0, 1e30)
    assert np.array_equal(quad, ref)
    idx = np.random.default_rng(4).choice(n, 1024, replace=False)
    a0 = _sample_accel(ctx, t.dy, t.dm, n, idx)
    e_mono, e_quad = _rms_rel(t.group(), a0, n, idx), _rms_rel(quad, a0, n, idx)
    assert e_quad[0] < 0.5 * e_mono[0], (e_quad, e_mono)
    net.close()
    t.close()


def test_quadrupole_walk_far_expansion_matches_oracle(ctx):
    """Quadrupole walk with far cells expanded about the group centre: the added
    term equals the oracle's per-body quadrupole of the near cells plus, for the
    far cells, the quadrupole field at the centre and its finite-difference
    Jacobian applied to each body's offset."""
    n, ratio, near_k = 4096, 1.6, 3.0
    y, m = _state("plummer", n, 26)
    t = Tree(ctx, y, m, ratio)
    a = t.arrays()
    body = _body_order(ctx, t)
    plain = t.group()
    df = ctx.alloc(6 * n)
    native.check(ctx.lib.nbh_bh_fcompute_group_quad(ctx.h, t.h, t.dy.ptr, df.ptr, near_k), "quad")
    got = df.read()
    c = t.counters()
    assert c[3] > 0 and c[4] > 0
    delta = (got - plain)[3 * n:].reshape(3, n).T
    want = np.zeros((n, 3))
    for g in range(n // 64):
        gn = n // 64 + g
        cen = 0.5 * (a["bmin"][gn] + a["bmax"][gn])
        # far/near split of every taken internal cell (rho_min 0: all far cells count)
        far, _, near = mo.group_band(n, g, a["cm"], a["crit"], a["bmin"], a["bmax"], 0.0, near_k)
        R = np.zeros(3)
        J = np.zeros((3, 3))
        for cell in far:
            l0, l1 = _leaf_slice(n, cell)
            _, Q = mo.second_moments(a["cm"][l0:l1], a["mass"][l0:l1])
            R += mo.quadrupole_field(cen, a["cm"][cell], Q)
            h = 1e-4 * np.linalg.norm(a["cm"][cell] - cen)
            for j in range(3):
                e = np.zeros(3); e[j] = h
                J[:, j] += (mo.quadrupole_field(cen + e, a["cm"][cell], Q) -
                            mo.quadrupole_field(cen - e, a["cm"][cell], Q)) / (2 * h)
        Qs = []
        for cell in near:
            l0, l1 = _leaf_slice(n, cell)
            Qs.append(mo.second_moments(a["cm"][l0:l1], a["mass"][l0:l1])[1])
        for leaf in range(n + 64 * g, n + 64 * g + 64):
            x = a["cm"][leaf]
            v = R + J @ (x - cen)
            for cell, Q in zip(near, Qs):
                v = v + mo.quadrupole_field(x, a["cm"][cell], Q)
            want[body[leaf]] = v
    err = np.linalg.norm(delta - want, axis=1) / np.abs(want).max()
    assert np.percentile(err, 99) < 1e-3 and err.max() < 5e-3, (np.percentile(err, 99), err.max())
    t.close()


def _moments(a, n, cells):
    return np.array([mo.second_moments(a["cm"][slice(*_leaf_slice(n, c))], a["mass"][slice(*_leaf_slice(n, c))])[1]
                     for c in cells]).reshape(-1, 6)


def _per_body(a, n, src, near, X):
    """Monopoles of the nodes src (r^2 clamped at 1e-8) and quadrupoles of the
    cells near at the points X."""
    r = a["cm"][src][None, :, :] - X[:, None, :]
    r2 = np.maximum(np.einsum("bci,bci->bc", r, r), 1e-8)
    v = np.einsum("bc,bci->bi", a["mass"][src][None, :] / (r2 * np.sqrt(r2)), r)
    return v + mo.quadrupole_fields(X, a["cm"][near], _moments(a, n, near))


def _far_expansion(a, n, far, cen, hd, X):
    """Far cells about cen: monopole to third order, quadrupole field plus its
    finite-difference Jacobian."""
    if not far:
        return np.zeros_like(X)
    f, J, D2, D3 = mo.local_expansion_tensors(a["cm"][far] - cen, a["mass"][far], 3)
    Qf = _moments(a, n, far)
    h = 1e-4 * hd
    gq = mo.quadrupole_fields(cen[None, :], a["cm"][far], Qf)[0]
    Jq = np.stack([(mo.quadrupole_fields((cen + h * e)[None, :], a["cm"][far], Qf)[0]
                    - mo.quadrupole_fields((cen - h * e)[None, :], a["cm"][far], Qf)[0]) / (2 * h)
                   for e in np.eye(3)], axis=1)
    return mo.local_expansion_eval(f + gq, J + Jq, D2, X - cen, D3)


def _group_terms(n, a, g, far, taken, near):
    gn = n // 64 + g
    cen = 0.5 * (a["bmin"][gn] + a["bmax"][gn])
    hd = 0.5 * np.linalg.norm(a["bmax"][gn] - a["bmin"][gn])
    X = a["cm"][n + 64 * g:n + 64 * g + 64]
    src = [t for t in taken if t >= n] + list(near)
    return X, _per_body(a, n, src, near, X) + _far_expansion(a, n, far, cen, hd, X)


def _oracle_local_walk(n, a, body, near_k):
    """Accelerations (original body order) of the quadrupole walk with a local far
    field: leaves and near cells per body (monopole, plus quadrupole for cells),
    far cells through the third-order expansion of their monopole and the
    first-order expansion of their quadrupole (finite-difference Jacobian) about
    the group centre."""
    want = np.zeros((n, 3))
    for g in range(n // 64):
        far, taken, near = mo.group_band(n, g, a["cm"], a["crit"], a["bmin"], a["bmax"], 0.0, near_k)
        _, v = _group_terms(n, a, g, far, taken, near)
        want[body[n + 64 * g:n + 64 * g + 64]] = v
    return want


def test_local_far_field_walk_matches_oracle(ctx):
    """nbh_bh_fcompute_group_local (far cells' monopole to third order and
    quadrupole to first order about the group centre, near cells and leaves per
    body) against the oracle's restatement of the same sums in FP64."""
    n, ratio, near_k = 4096, 1.6, 3.0
    y, m = _state("plummer", n, 27)
    t = Tree(ctx, y, m, ratio)
    a = t.arrays()
    df = ctx.alloc(6 * n)
    native.check(ctx.lib.nbh_bh_fcompute_group_local(ctx.h, t.h, t.dy.ptr, df.ptr, near_k), "local")
    got = df.read()
    c = t.counters()
    assert c[2] == 0 and c[3] > 0 and c[4] > 0
    assert np.array_equal(got[:3 * n], y[3 * n:])
    want = _oracle_local_walk(n, a, _body_order(ctx, t), near_k)
    acc = got[3 * n:].reshape(3, n).T
    err = np.linalg.norm(acc - want, axis=1) / np.linalg.norm(want, axis=1)
    print(f"\n[local walk vs oracle] rel err p50 {np.median(err):.2e} p99 {np.percentile(err, 99):.2e} "
          f"max {err.max():.2e}")
    assert np.percentile(err, 99) < 1e-4 and err.max() < 3e-3, (np.percentile(err, 99), err.max())
    bad = ctx.alloc(6 * n)
    assert ctx.lib.nbh_bh_fcompute_group_local(ctx.h, t.h, t.dy.ptr, bad.ptr, 0.5) != 0
    t.close()


def test_two_level_local_walk_matches_oracle(ctx):
    """nbh_bh_fcompute_group_local2 on the bodies of two super-groups: the
    oracle's super-group walk (far cells expanded about the super-group's box
    centre, the rest listed) and each group's walk from the list (near cells and
    leaves per body, farther cells expanded about the group centre)."""
    n, ratio, near_k = 1 << 17, 1.0 / 0.6, 3.0
    y, m = _state("plummer", n, 29)
    t = Tree(ctx, y, m, ratio)
    a = t.arrays()
    body = _body_order(ctx, t)
    df = ctx.alloc(6 * n)
    native.check(ctx.lib.nbh_bh_fcompute_group_local2(ctx.h, t.h, t.dy.ptr, df.ptr, near_k), "local2")
    got = df.read()[3 * n:].reshape(3, n).T
    assert t.counters()[2] == 0                     # no list overflow, no spill
    errs, shared = [], 0
    for s in (5, 100):
        sg = (n >> 10) + s
        scen = 0.5 * (a["bmin"][sg] + a["bmax"][sg])
        shd = 0.5 * np.linalg.norm(a["bmax"][sg] - a["bmin"][sg])
        # the super-group level uses max(near_k, 5) super-group half-diagonals
        far_s, listed = mo.super_band(n, s, a["cm"], a["crit"], a["bmin"], a["bmax"], max(near_k, 5.0), ratio,
                                      shift=4)
        shared += len(far_s)
        for gsub in range(0, 16, 2):
            g = 16 * s + gsub
            far, taken, near = mo.group_band_from(n, g, listed, a["cm"], a["crit"], a["bmin"], a["bmax"], near_k)
            X, v = _group_terms(n, a, g, far, taken, near)
            v = v + _far_expansion(a, n, far_s, scen, shd, X)
            b = body[n + 64 * g:n + 64 * g + 64]
            errs.append(np.linalg.norm(got[b] - v, axis=1) / np.linalg.norm(v, axis=1))
    err = np.concatenate(errs)
    print(f"\n[two-level walk vs oracle] shared far cells {shared}, rel err p50 {np.median(err):.2e} "
          f"p99 {np.percentile(err, 99):.2e} max {err.max():.2e}")
    assert shared > 0
    assert np.percentile(err, 99) < 1e-4 and err.max() < 3e-3, (np.percentile(err, 99), err.max())
    t.close()


def test_two_level_walk_overflow_fallback(ctx):
    """At 2^22 Plummer bodies with the default list capacity no super-group
    falls back; with a capacity of 2048 entries some lists overflow and their
    groups walk from the root instead.  Both keep the one-level local walk's
    accuracy against the FP64 direct sum."""
    n, ratio, near_k = 1 << 22, 1.0 / 0.6, 3.0
    y, m = _state("plummer", n, 4)
    t = Tree(ctx, y, m, ratio)
    idx = np.random.default_rng(6).choice(n, 1024, replace=False)
    a0 = _sample_accel(ctx, t.dy, t.dm, n, idx)
    df = ctx.alloc(6 * n)
    native.check(ctx.lib.nbh_bh_fcompute_group_local(ctx.h, t.h, t.dy.ptr, df.ptr, near_k), "local")
    e1 = _rms_rel(df.read(), a0, n, idx)
    native.check(ctx.lib.nbh_bh_fcompute_group_local2(ctx.h, t.h, t.dy.ptr, df.ptr, near_k), "local2")
    full = t.counters()[2]
    e2 = _rms_rel(df.read(), a0, n, idx)
    native.check(ctx.lib.nbh_tree_set_list_cap(ctx.h, t.h, 2048), "cap")
    native.check(ctx.lib.nbh_bh_fcompute_group_local2(ctx.h, t.h, t.dy.ptr, df.ptr, near_k), "local2, cap 2048")
    overflowed = t.counters()[2]
    e3 = _rms_rel(df.read(), a0, n, idx)
    print(f"\n[two-level walk, 2^22] fell back: {full} (cap 32768), {overflowed} of {n >> 10} (cap 2048); "
          f"rms one-level {e1[0]:.3e}, two-level {e2[0]:.3e}, cap 2048 {e3[0]:.3e}")
    assert full == 0 and overflowed > 0
    assert e2[0] < 1.1 * e1[0] and e3[0] < 1.1 * e1[0] and e3[0] < 1.2e-3
    assert ctx.lib.nbh_tree_set_list_cap(ctx.h, t.h, 8) != 0      # below the minimum: rejected
    t.close()


def _tree_bytes(ctx, t):
    b = C.c_size_t()
    native.check(ctx.lib.nbh_tree_bytes(ctx.h, t.h, C.byref(b)), "tree_bytes")
    return b.value


def test_two_level_list_allocation_is_all_or_nothing(ctx, monkeypatch):
    """A failed allocation of the two-level walk's lists (forced through the
    library's NBH_TEST_FAIL_ALLOC hook at each of its three allocations, in
    nbh_tree_set_list_cap and in the walk's own first-use allocation) leaves
    none of them allocated and launches nothing; the next call allocates them
    afresh and the walk equals that of a tree that never failed."""
    n, ratio, near_k, cap = 1 << 14, 1.0 / 0.6, 3.0, 4096
    y, m = _state("plummer", n, 31)
    t = Tree(ctx, y, m, ratio)
    base = _tree_bytes(ctx, t)
    lists = (n >> 10) * (cap * 4 + 4 + 34 * 4)
    df = ctx.alloc(6 * n)
    for k in (1, 2, 3):
        monkeypatch.setenv("NBH_TEST_FAIL_ALLOC", str(k))
        assert ctx.lib.nbh_tree_set_list_cap(ctx.h, t.h, cap) == 3          # NBH_ERR_NOMEM
        assert _tree_bytes(ctx, t) == base
        assert ctx.lib.nbh_bh_fcompute_group_local2(ctx.h, t.h, t.dy.ptr, df.ptr, near_k) == 3
        assert _tree_bytes(ctx, t) == base
        monkeypatch.delenv("NBH_TEST_FAIL_ALLOC")
    native.check(ctx.lib.nbh_bh_fcompute_group_local2(ctx.h, t.h, t.dy.ptr, df.ptr, near_k), "local2")
    assert _tree_bytes(ctx, t) == base + lists
    got = df.read()
    t2 = Tree(ctx, y, m, ratio)
    native.check(ctx.lib.nbh_tree_set_list_cap(ctx.h, t2.h, cap), "cap")
    assert _tree_bytes(ctx, t2) == base + lists
    native.check(ctx.lib.nbh_bh_fcompute_group_local2(ctx.h, t2.h, t2.dy.ptr, df.ptr, near_k), "local2")
    assert np.array_equal(got, df.read())
    t2.close()
    t.close()


def test_local_far_field_walk_accuracy(ctx):
    """At theta 0.6 the local far field keeps the quadrupole walk's accuracy
    against the FP64 direct sum (every cell per body as the reference point)."""
    n, ratio, near_k = 1 << 16, 1.0 / 0.6, 3.0
    y, m = _state("plummer", n, 28)
    t = Tree(ctx, y, m, ratio)
    idx = np.random.default_rng(5).choice(n, 2048, replace=False)
    a0 = _sample_accel(ctx, t.dy, t.dm, n, idx)
    df = ctx.alloc(6 * n)
    native.check(ctx.lib.nbh_bh_fcompute_group_quad(ctx.h, t.h, t.dy.ptr, df.ptr, float("inf")), "quad")
    e_all = _rms_rel(df.read(), a0, n, idx)
    native.check(ctx.lib.nbh_bh_fcompute_group_local(ctx.h, t.h, t.dy.ptr, df.ptr, near_k), "local")
    e_local = _rms_rel(df.read(), a0, n, idx)
    native.check(ctx.lib.nbh_bh_fcompute_group_local2(ctx.h, t.h, t.dy.ptr, df.ptr, near_k), "local2")
    e_local2 = _rms_rel(df.read(), a0, n, idx)
    e_mono = _rms_rel(t.group(), a0, n, idx)
    print(f"\n[local walk accuracy] rms/p99: every cell per body {e_all[0]:.3e}/{e_all[1]:.3e}  "
          f"local far field {e_local[0]:.3e}/{e_local[1]:.3e}  two-level {e_local2[0]:.3e}/{e_local2[1]:.3e}  "
          f"monopole {e_mono[0]:.3e}/{e_mono[1]:.3e}")
    assert e_local[0] < 1.15 * e_all[0] and e_local[0] < 0.5 * e_mono[0]
    assert e_local2[0] < 1.15 * e_all[0]
    t.close()


def test_corrected_walk_matches_oracle(ctx):
    """Correction added by the kernel == the oracle's near-cell quadrupoles plus
    its network evaluated on the oracle's far cells, expanded to first order
    about each group centre."""
    n, ratio, rho, near_k = 4096, 1.25, 0.1, 2.0
    y, m = _state("plummer", n, 23)
    t = Tree(ctx, y, m, ratio)
    a = t.arrays()
    net = Net(ctx)
    p = _random_params(31, bias=0.1) * F32(0.5)
    net.set(p)
    plain = t.group()
    got = t.group_mlp(net, rho, near_k)
    c = t.counters()
    assert c[3] > 0 and c[4] > 0
    delta = (got - plain)[3 * n:].reshape(3, n).T
    want = _oracle_correction(n, a, _body_order(ctx, t), p, rho, near_k)
    scale = np.abs(want).max()
    assert scale > 0
    err = np.abs(delta - want).max() / scale
    assert err < 2e-2, err
    net.close()
    t.close()


def _sample_accel(ctx, dy, dm, n, idx):
    di = ctx.put_raw(np.ascontiguousarray(idx, np.uint32))
    da = ctx.alloc(3 * len(idx))
    native.check(ctx.lib.nbh_direct_sample(ctx.h, dy.ptr, dm.ptr, n, di.ptr, len(idx), da.ptr), "sample")
    return da.read().reshape(-1, 3)


def _rms_rel(f, a0, n, idx):
    a = f[3 * n:].reshape(3, n).T[idx]
    e = np.linalg.norm(a - a0, axis=1) / np.linalg.norm(a0, axis=1)
    return float(np.sqrt(np.mean(e ** 2))), float(np.percentile(e, 99))


def test_trained_correction_gain(ctx, capsys):
    """Train on the far cells of one Plummer sphere, apply to another: at the
    same opening ratio, the trained network lowers the error of the walk with
    the near-cell quadrupoles alone, and both beat the plain walk."""
    n, ratio, rho, near_k = 1 << 16, 1.25, 0.0, 2.0
    ytr, mtr = _state("plummer", n, 31)
    ttr = Tree(ctx, ytr, mtr, ratio)
    groups = np.arange(n // 64)
    cap = 1 << 22
    smp = ttr.samples(groups, rho, near_k, cap)
    total = smp.total
    k = min(total, cap)
    net = Net(ctx, 5)
    el = net.train(smp.dx, smp.db, smp.dt, k, 30, 1024, 2e-3)
    ttr.close()
    y, m = _state("plummer", n, 32)
    t = Tree(ctx, y, m, ratio)
    idx = np.random.default_rng(3).choice(n, 4096, replace=False)
    a0 = _sample_accel(ctx, t.dy, t.dm, n, idx)
    e_plain = _rms_rel(t.group(), a0, n, idx)
    zero = Net(ctx)
    zero.set(np.zeros(mo.N_PARAMS, F32))
    e_near = _rms_rel(t.group_mlp(zero, rho, near_k), a0, n, idx)
    e_full = _rms_rel(t.group_mlp(net, rho, near_k), a0, n, idx)
    t2 = Tree(ctx, y, m, 2.0)
    e_plain2 = _rms_rel(t2.group(), a0, n, idx)
    with capsys.disabled():
        print(f"\n[correction gain] samples {total}  loss {el[0]:.3g}->{el[-1]:.3g}  rms/p99: "
              f"plain 1.25 {e_plain[0]:.3e}/{e_plain[1]:.3e}  near-quad 1.25 {e_near[0]:.3e}/{e_near[1]:.3e}  "
              f"near-quad+net 1.25 {e_full[0]:.3e}/{e_full[1]:.3e}  plain 2.0 {e_plain2[0]:.3e}/{e_plain2[1]:.3e}")
    # measured on the MI355X: plain 1.40e-2, near-quad 1.05e-2, near-quad + net 8.1e-3
    assert e_near[0] < 0.9 * e_plain[0]
    assert e_full[0] < 0.9 * e_near[0]
    net.close()
    zero.close()
    t.close()
    t2.close()


def _super_box(a, n, s):
    sg = (n >> 10) + s
    return 0.5 * (a["bmin"][sg] + a["bmax"][sg]), sg


def test_super_samples_match_oracle(ctx):
    """nbh_bh_super_samples: the rows of a super-group are exactly the far cells
    of the oracle's super-group walk (mo.super_band), and each row's features,
    basis and target (field remainder after the monopole and the quadrupole at
    the super box centre, by brute force over the cell's bodies) match the
    oracle's.  Most super-groups of a Plummer sphere this small have boxes that
    reach the outskirts, so that every cell lies within 5 of their
    half-diagonals and none is far; the super-groups are chosen from the
    oracle's walk: three with far cells and one without."""
    n, ratio, near_k = 1 << 16, 1.25, 5.0
    y, m = _state("plummer", n, 41)
    t = Tree(ctx, y, m, ratio)
    a = t.arrays()
    far_of = {s: mo.super_band(n, s, a["cm"], a["crit"], a["bmin"], a["bmax"], max(near_k, 5.0), ratio, shift=4)[0]
              for s in range(n >> 10)}
    with_far = [s for s in far_of if far_of[s]]
    without = [s for s in far_of if not far_of[s]]
    assert len(with_far) >= 3 and without
    supers = with_far[:3] + without[:1]
    smp = t.super_samples(near_k, supers, 1 << 16)
    assert smp.total == len(smp.cells) == sum(len(far_of[s]) for s in supers)
    for s in supers:
        far = far_of[s]
        got = [int(c) for ss, c in smp.cells if ss == s]
        assert got == [int(c) for c in far] or sorted(got) == sorted(int(c) for c in far)
    rows = np.random.default_rng(1).choice(len(smp.cells), min(200, len(smp.cells)), replace=False)
    for k in rows:
        s, c = int(smp.cells[k, 0]), int(smp.cells[k, 1])
        p, _ = _super_box(a, n, s)
        l0, l1 = _leaf_slice(n, c)
        fx, fb, ft = mo.target_vector(p, a["cm"][l0:l1], a["mass"][l0:l1], a["bmin"][c], a["bmax"][c], order=2)
        assert np.allclose(smp.x[k], fx, rtol=1e-5, atol=1e-6 * np.abs(fx).max()), (k, smp.x[k], fx)
        assert np.allclose(smp.b[k], fb, rtol=1e-5, atol=1e-6 * np.abs(fb).max()), (k, smp.b[k], fb)
        assert np.allclose(smp.t[k], ft, rtol=1e-3, atol=1e-4 * np.abs(ft).max()), (k, smp.t[k], ft)
    t.close()


def test_learned_two_level_walk_matches_oracle(ctx):
    """nbh_bh_fcompute_group_local2_mlp minus nbh_bh_fcompute_group_local2 is
    the network's correction: for every body of a super-group, the sum over the
    oracle's far cells of that super-group of R + J (x - p), with R and J the
    network's predictions (BF16 emulation) scaled back.  An all-zero network
    changes nothing, bit for bit."""
    n, ratio, near_k = 1 << 16, 1.25, 5.0
    y, m = _state("plummer", n, 42)
    t = Tree(ctx, y, m, ratio)
    a = t.arrays()
    body = _body_order(ctx, t)
    # super-groups with far cells (the others get no correction: checked below)
    far_of = {s: mo.super_band(n, s, a["cm"], a["crit"], a["bmin"], a["bmax"], max(near_k, 5.0), ratio, shift=4)[0]
              for s in range(n >> 10)}
    plain = t.local2(near_k)
    zero = Net(ctx)
    zero.set(np.zeros(mo.N_PARAMS, F32))
    assert np.array_equal(t.local2(near_k, zero), plain)
    net = Net(ctx, 9)
    p = _random_params(13, 0.1)
    net.set(p)
    got = t.local2(near_k, net)
    diff = (got - plain)[3 * n:].reshape(3, n).T
    checked = 0
    for s in [s for s in far_of if not far_of[s]][:4]:
        leaves = body[n + 1024 * s:n + 1024 * s + 1024]
        assert np.all(diff[leaves] == 0.0), s
    for s in [s for s in far_of if far_of[s]][:3]:
        far = far_of[s]
        c0, sg = _super_box(a, n, s)
        X, B, W0, W1 = [], [], [], []
        for c in far:
            l0, l1 = _leaf_slice(n, c)
            Q = mo.second_moments(a["cm"][l0:l1], a["mass"][l0:l1])[1]
            f, b, scale, d = mo.cell_features(c0, a["cm"][c], a["mass"][c], a["bmin"][c], a["bmax"][c], Q)
            X.append(f); B.append(b); W0.append(scale * f[0]); W1.append(scale * f[0] / (mo.K_JAC * d))
        pred = mo.predict(p, np.array(X, np.float32), np.array(B, np.float32), "bf16").astype(np.float64)
        R = (pred[:, :3] * np.array(W0)[:, None]).sum(0)
        Jv = (pred[:, 3:] * np.array(W1)[:, None]).sum(0)
        J = np.array([[Jv[0], Jv[3], Jv[4]], [Jv[3], Jv[1], Jv[5]], [Jv[4], Jv[5], Jv[2]]])
        leaves = body[n + 1024 * s:n + 1024 * s + 1024]
        X0 = a["cm"][n + 1024 * s:n + 1024 * s + 1024]
        want = R[None, :] + (X0 - c0) @ J.T
        err = np.linalg.norm(diff[leaves] - want, axis=1) / np.linalg.norm(want, axis=1)
        assert np.percentile(err, 99) < 2e-3 and err.max() < 1e-2, (s, np.percentile(err, 99), err.max())
        checked += 1
    assert checked == 3
    zero.close()
    net.close()
    t.close()


def test_learned_two_level_gain(ctx, capsys):
    """Train the two-level walk's correction on the super-group far cells of one
    Plummer sphere (theta 0.8, near 5) and apply it to another.  What the
    network corrects -- the remainder after the monopole and quadrupole terms
    of a super-group's far cells, summed over them at the super box centre, in
    physical units -- falls by more than half on the held-out sphere for the
    median super-group, and the held-out loss is below that of predicting zero.
    The walk's whole error is reported beside it: at this size that remainder is
    a small part of it (measured: RMS 3.81e-3 without, 3.84e-3 with the
    network), so the test only holds the corrected walk within 2% of it."""
    n, ratio, near_k = 1 << 17, 1.25, 5.0
    ytr, mtr = _state("plummer", n, 43)
    ttr = Tree(ctx, ytr, mtr, ratio)
    smp = ttr.super_samples(near_k, np.arange(n >> 10), 1 << 22)
    k = min(smp.total, 1 << 22)
    net = Net(ctx, 5)
    el = net.train(smp.dx, smp.db, smp.dt, k, 40, 1024, 2e-3)
    ttr.close()
    y, m = _state("plummer", n, 44)
    t = Tree(ctx, y, m, ratio)
    a = t.arrays()
    hold = t.super_samples(near_k, np.arange(n >> 10), 1 << 22)
    kh = min(hold.total, 1 << 22)
    assert kh == hold.total > 0
    _, loss_net = net.gradient(hold.dx, hold.db, hold.dt, kh)
    loss_zero = float(np.mean(np.sum(hold.t.astype(np.float64) ** 2, axis=1)))
    # remainder of each super-group's far cells at its box centre: target and
    # prediction scaled back by the row's weight (cell_features scale x rho)
    pred = mo.predict(net.get(), hold.x, hold.b, "bf16").astype(np.float64)
    rows = {}
    for r, (sg, c) in enumerate(hold.cells):
        rows.setdefault(int(sg), []).append(r)
    ratios = []
    for sg, rr in rows.items():
        c0, _ = _super_box(a, n, sg)
        w = np.empty(len(rr))
        for q, r in enumerate(rr):
            c = int(hold.cells[r, 1])
            l0, l1 = _leaf_slice(n, c)
            Q = mo.second_moments(a["cm"][l0:l1], a["mass"][l0:l1])[1]
            f, _, scale, _ = mo.cell_features(c0, a["cm"][c], a["mass"][c], a["bmin"][c], a["bmax"][c], Q)
            w[q] = scale * f[0]
        rt = (hold.t[rr, :3].astype(np.float64) * w[:, None]).sum(0)
        rp = (pred[rr, :3] * w[:, None]).sum(0)
        ratios.append(np.linalg.norm(rt - rp) / np.linalg.norm(rt))
    idx = np.random.default_rng(3).choice(n, 4096, replace=False)
    a0 = _sample_accel(ctx, t.dy, t.dm, n, idx)
    e_plain = _rms_rel(t.local2(near_k), a0, n, idx)
    e_net = _rms_rel(t.local2(near_k, net), a0, n, idx)
    with capsys.disabled():
        print(f"\n[two-level correction gain] train rows {smp.total}, loss {el[0]:.3g}->{el[-1]:.3g}; held-out "
              f"loss {loss_net:.3g} (zero prediction {loss_zero:.3g}); super-group far-field remainder left by "
              f"the network (median over {len(ratios)} super-groups) {np.median(ratios):.3f}; rms/p99 two-level "
              f"theta 0.8: {e_plain[0]:.4e}/{e_plain[1]:.4e} -> with the network {e_net[0]:.4e}/{e_net[1]:.4e}")
    assert loss_net < 0.8 * loss_zero
    assert np.median(ratios) < 0.5
    assert e_net[0] < 1.02 * e_plain[0]
    net.close()
    t.close()


def test_two_level_walk_stack_overflow_falls_back(ctx, monkeypatch):
    """When the super-group walk runs out of stack (forced here through the
    library's NBH_TEST_SUPER_STACK hook, 64 entries), every group of that
    super-group ignores the partial list and expansion and walks from the root:
    its bodies get the one-level local walk's result bit for bit, the bodies of
    super-groups that did not run out get the normal two-level result bit for
    bit, and the fallbacks are counted in walk counter out[2].  (At this size
    most super-groups list the root at once and never push: measured, 9 of 64
    run out.)"""
    n, ratio, near_k = 1 << 16, 1.0 / 0.6, 3.0
    y, m = _state("plummer", n, 45)
    t = Tree(ctx, y, m, ratio)
    df = ctx.alloc(6 * n)
    native.check(ctx.lib.nbh_bh_fcompute_group_local(ctx.h, t.h, t.dy.ptr, df.ptr, near_k), "local")
    one = df.read()
    monkeypatch.setenv("NBH_TEST_SUPER_STACK", "64")
    two = t.local2(near_k)
    c = t.counters()
    monkeypatch.delenv("NBH_TEST_SUPER_STACK")
    normal = t.local2(near_k)
    assert t.counters()[2] == 0 and not np.array_equal(normal, one)
    body = _body_order(ctx, t)
    fell = 0
    for s in range(n >> 10):
        leaves = body[n + 1024 * s:n + 1024 * (s + 1)].astype(np.int64)
        rows = np.concatenate([leaves + k * n for k in range(6)])
        if np.array_equal(two[rows], one[rows]) and not np.array_equal(normal[rows], one[rows]):
            fell += 1
        else:
            assert np.array_equal(two[rows], normal[rows]), s
    assert 1 <= fell <= c[2] <= n >> 10, (fell, c[2])
    t.close()
