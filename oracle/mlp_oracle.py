"""CPU oracle of the Barnes-Hut correction network and of its training data.

TEST INFRASTRUCTURE ONLY: imported by tests/ (and nowhere in the product path).
The network has no counterpart in the reference; what it approximates does:
the field of a tree cell's bodies (force law of nbody/nbody_engine_simple.cpp:
26-55, r^2 clamped at 1e-8) minus the monopole the heap walk uses in its place
(nbody/nbody_space_heap.cpp:94-133, nbody_space_heap_stackless.cpp:3-26).

Contents
  * unpack / forward / loss_grad: the 16-64-64-9 tanh MLP on a flat parameter
    vector (W1[16][64] b1[64] W2[64][64] b2[64] W3[64][16] b3[16], row-major),
    in float64, or emulating the device arithmetic (every matrix-core operand
    rounded to BF16 with round-to-nearest-even, FP32 elsewhere);
  * cell_features / basis_matrix / predict: the 16 rotation-invariant inputs of
    (evaluation point, cell) and the basis the 9 outputs are coefficients of;
  * cell_residual: exact residual field and Jacobian of a cell by brute force;
    cell_remainder: the same after the quadrupole term too (the target of the
    two-level walk's correction), with quadrupole_jacobian;
  * group_band: the near and far cells of one 64-body group, by a plain stack
    walk of the heap tree with the group walk's acceptance rule; super_band /
    group_band_from: the two walks of the two-level local walk (super-group of
    2^shift groups, 16 in the kernels, then each group from the super-group's list);
  * quadrupole_field: the textbook quadrupole term of a cell's field;
  * local_expansion_tensors / local_expansion_eval: second- and third-order
    Taylor expansions of cells' monopole fields about a group centre (the far
    field of nbh_bh_fcompute_group_local), and quadrupole_fields (vectorised
    quadrupole_field summed over cells).

Pinning (tests/test_oracle_mlp.py)
  * loss_grad (float64): central finite differences of the loss, every
    parameter block;  bf16 emulation with rounding disabled equals float64;
  * bf16 rounding: exhaustive on special patterns, ties to even, vs float32;
  * cell_residual: Jacobian vs finite differences of the field; at small
    size/distance the residual tends to the textbook quadrupole field;
  * cell_remainder: shrinking a cell about its mass centre by 2 divides the
    remainder by about 8 (octupole order), where a dropped or doubled
    quadrupole term would divide it by 4;
  * cell_features / basis: invariance of the features and covariance of the
    basis under a symmetry of the box; the quadrupole coefficients reproduce
    the textbook quadrupole field and its finite-difference Jacobian;
  * group_band: a cell set property (every body of the group is covered once by
    accepted cells and opened cells' descendants), checked on small trees;
  * quadrupole_field: limit of the exact residual of a small distant cell;
  * local_expansion_*: field at the centre equals the exact monopole field,
    first derivatives equal finite differences, and the error of the expansion
    falls as |delta|^3 at second order and |delta|^4 at third (halving the offset
    divides it by about 8 and 16, which a wrong coefficient of any order would
    not).
"""
from __future__ import annotations

import numpy as np

K_IN, K_HID, K_OUTP, K_OUT = 16, 64, 16, 9
OFF_W1 = 0
OFF_B1 = OFF_W1 + K_IN * K_HID
OFF_W2 = OFF_B1 + K_HID
OFF_B2 = OFF_W2 + K_HID * K_HID
OFF_W3 = OFF_B2 + K_HID
OFF_B3 = OFF_W3 + K_HID * K_OUTP
N_PARAMS = OFF_B3 + K_OUTP  # 6288
K_JAC = 0.25                # Jacobian outputs are K_JAC * J * d / scale


def bf16(a) -> np.ndarray:
    """Round float32 values to the nearest BF16 (ties to even), as float32."""
    a = np.ascontiguousarray(a, dtype=np.float32)
    u = a.view(np.uint32).astype(np.uint64)
    u = (u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000
    return u.astype(np.uint32).view(np.float32).reshape(a.shape)


def unpack(p):
    p = np.asarray(p)
    assert p.size == N_PARAMS
    W1 = p[OFF_W1:OFF_B1].reshape(K_IN, K_HID)
    b1 = p[OFF_B1:OFF_W2]
    W2 = p[OFF_W2:OFF_B2].reshape(K_HID, K_HID)
    b2 = p[OFF_B2:OFF_W3]
    W3 = p[OFF_W3:OFF_B3].reshape(K_HID, K_OUTP)
    b3 = p[OFF_B3:]
    return W1, b1, W2, b2, W3, b3


def glorot(seed: int) -> np.ndarray:
    """Glorot-uniform weights and zero biases (output columns >= 9 zero)."""
    rng = np.random.default_rng(seed)
    p = np.zeros(N_PARAMS, np.float32)
    for off, fi, fo, used in ((OFF_W1, K_IN, K_HID, K_HID), (OFF_W2, K_HID, K_HID, K_HID),
                              (OFF_W3, K_HID, K_OUTP, K_OUT)):
        a = np.sqrt(6.0 / (fi + used))
        w = rng.uniform(-a, a, size=(fi, fo))
        w[:, used:] = 0.0
        p[off:off + fi * fo] = w.ravel()
    return p


def _ops(mode: str):
    if mode == "f64":
        return np.float64, (lambda a: np.asarray(a, np.float64))
    if mode == "bf16":
        return np.float32, bf16
    if mode == "f32":   # bf16 path with rounding disabled
        return np.float32, (lambda a: np.asarray(a, np.float32))
    raise ValueError(mode)


def forward(p, x, mode: str = "f64"):
    """Outputs [m][9] (and the activations) of the network on features x[m][16]."""
    dt, rnd = _ops(mode)
    W1, b1, W2, b2, W3, b3 = (np.asarray(v, dt) for v in unpack(p))
    x = np.asarray(x, dt)
    xb = rnd(x)
    h1 = np.tanh(rnd(xb) @ rnd(W1) + b1).astype(dt)
    h2 = np.tanh(rnd(h1) @ rnd(W2) + b2).astype(dt)
    out = (rnd(h2) @ rnd(W3) + b3).astype(dt)
    return out[:, :K_OUT], (xb, h1, h2, out)


def basis_matrix(b) -> np.ndarray:
    """[m][9 columns][9 components]: the target-space vector of each output
    column (R: 3 components, then the symmetric Jacobian xx yy zz xy xz yz)."""
    b = np.asarray(b)
    m = b.shape[0]
    u, w, o, q = b[:, 0:3], b[:, 3:6], b[:, 6:9], b[:, 9:15]

    def sym(a, c):
        return np.stack([2 * a[:, 0] * c[:, 0], 2 * a[:, 1] * c[:, 1], 2 * a[:, 2] * c[:, 2],
                         a[:, 0] * c[:, 1] + a[:, 1] * c[:, 0], a[:, 0] * c[:, 2] + a[:, 2] * c[:, 0],
                         a[:, 1] * c[:, 2] + a[:, 2] * c[:, 1]], axis=1)

    B = np.zeros((m, K_OUT, K_OUT), dtype=b.dtype)
    B[:, 0, :3], B[:, 1, :3], B[:, 2, :3] = u, w, o
    B[:, 3, 3:] = q
    B[:, 4, 3:] = sym(u, w)
    B[:, 5, 3:6] = 1.0
    B[:, 6, 3:] = 0.5 * sym(u, u)
    B[:, 7, 3:] = sym(u, o)
    B[:, 8, 3:] = 0.5 * sym(o, o)
    return B


def predict(p, x, b, mode: str = "f64"):
    """Predicted targets [m][9] = sum over columns of coefficient x basis vector."""
    c, _ = forward(p, x, mode)
    dt, _ = _ops(mode)
    return np.einsum("mc,mck->mk", c, basis_matrix(np.asarray(b, dt))).astype(dt)


def loss_grad(p, x, b, t, mode: str = "f64"):
    """L = sum |B c - t|^2 / m and dL/dp (flat), matching nbh_mlp_gradient."""
    dt, rnd = _ops(mode)
    W1, b1, W2, b2, W3, b3 = (np.asarray(v, dt) for v in unpack(p))
    m = x.shape[0]
    _, (xb, h1, h2, out) = forward(p, x, mode)
    B = basis_matrix(np.asarray(b, dt))
    pred = np.einsum("mc,mck->mk", out[:, :K_OUT], B)
    diff = (pred - np.asarray(t, dt)).astype(dt)
    loss = float(np.sum(diff.astype(np.float64) ** 2) / m)
    dout = np.zeros_like(out)
    dout[:, :K_OUT] = 2.0 * np.einsum("mk,mck->mc", diff, B) / m
    gW3 = rnd(h2).T @ rnd(dout)
    gb3 = dout.sum(0)
    dz2 = (rnd(dout) @ rnd(W3).T) * (1 - h2 * h2)
    gW2 = rnd(h1).T @ rnd(dz2)
    gb2 = dz2.sum(0)
    dz1 = (rnd(dz2) @ rnd(W2).T) * (1 - h1 * h1)
    gW1 = rnd(xb).T @ rnd(dz1)
    gb1 = dz1.sum(0)
    g = np.concatenate([gW1.ravel(), gb1, gW2.ravel(), gb2, gW3.ravel(), gb3]).astype(np.float64)
    return loss, g


# ---------------------------------------------------------------------------
# physics side
# ---------------------------------------------------------------------------
def cell_radius(bmin, bmax, cm) -> float:
    """Bounding radius of a cell as in the opening test (nbody_space_heap.cpp:130-133)."""
    bmin, bmax, cm = (np.asarray(v, np.float64) for v in (bmin, bmax, cm))
    return 0.5 * np.linalg.norm(bmax - bmin) + np.linalg.norm(0.5 * (bmax + bmin) - cm)


def cell_features(p, cm, M, bmin, bmax, Q):
    """16 rotation-invariant features, the basis data b[16] (uh, w, o, q, 0),
    the target scale M r^2 / d^4 and the distance d of (point p, cell)."""
    p, cm, Q = (np.asarray(v, np.float64) for v in (p, cm, Q))
    bmin, bmax = np.asarray(bmin, np.float64), np.asarray(bmax, np.float64)
    u = cm - p
    d = np.linalg.norm(u)
    r = cell_radius(bmin, bmax, cm)
    q = np.array([[Q[0], Q[3], Q[4]], [Q[3], Q[1], Q[5]], [Q[4], Q[5], Q[2]]]) / (M * r * r)
    h = u / d
    w = q @ h
    o = (0.5 * (bmax + bmin) - cm) / r
    v = q @ o
    rho = r / d
    uqu = h @ w
    uo = h @ o
    f = np.array([rho, rho * rho, np.trace(q), uqu, w @ w, np.sum(q * q), uo, o @ o, w @ o, o @ v,
                  rho * uqu, rho * np.trace(q), uo * uo, rho * uo, v @ v, uqu * uqu])
    b = np.concatenate([h, w, o, [q[0, 0], q[1, 1], q[2, 2], q[0, 1], q[0, 2], q[1, 2]], [0.0]])
    return f, b, M * rho * rho / (d * d), d


def field(p, pos, mass):
    """Field sum_j m_j (x_j - p) / |x_j - p|^3 (r^2 >= 1e-8) and its Jacobian."""
    r = np.asarray(pos, np.float64) - np.asarray(p, np.float64)
    r2 = np.maximum(np.einsum("ij,ij->i", r, r), 1e-8)
    w = mass / (r2 * np.sqrt(r2))
    a = (r * w[:, None]).sum(0)
    J = (3.0 * (w / r2)[:, None, None] * r[:, :, None] * r[:, None, :]).sum(0) - np.eye(3) * w.sum()
    return a, J


def second_moments(pos, mass):
    cm = (pos * mass[:, None]).sum(0) / mass.sum()
    d = pos - cm
    Q = np.einsum("i,ij,ik->jk", mass, d, d)
    return cm, np.array([Q[0, 0], Q[1, 1], Q[2, 2], Q[0, 1], Q[0, 2], Q[1, 2]])


def cell_residual(p, pos, mass):
    """Exact field of the bodies minus their monopole at p, and its Jacobian."""
    a, J = field(p, pos, mass)
    M = mass.sum()
    cm = (pos * mass[:, None]).sum(0) / M
    a0, J0 = field(p, cm[None, :], np.array([M]))
    return a - a0, J - J0


def quadrupole_jacobian(p, cm, Q, h=None):
    """d quadrupole_field / dp by central differences of the textbook field
    (step 1e-4 of the distance: truncation error ~1e-8 relative)."""
    p = np.asarray(p, np.float64)
    if h is None:
        h = 1e-4 * np.linalg.norm(np.asarray(cm, np.float64) - p)
    cols = [(quadrupole_field(p + h * e, cm, Q) - quadrupole_field(p - h * e, cm, Q)) / (2 * h) for e in np.eye(3)]
    return np.stack(cols, 1)


def cell_remainder(p, pos, mass):
    """Exact field of the bodies minus their monopole and quadrupole terms at
    p (the octupole and higher terms), and its Jacobian."""
    R, J = cell_residual(p, pos, mass)
    cm, Q = second_moments(pos, mass)
    return R - quadrupole_field(p, cm, Q), J - quadrupole_jacobian(p, cm, Q)


def target_vector(p, pos, mass, bmin, bmax, order: int = 1):
    """Features x[16], basis b[16] and target t[9] of the cell holding bodies
    pos/mass: order 1, the residual after the monopole, scaled by
    M rho^2 / d^2; order 2, the remainder after the monopole and the
    quadrupole, scaled by M rho^3 / d^2."""
    cm, Q = second_moments(pos, mass)
    M = mass.sum()
    f, b, scale, d = cell_features(p, cm, M, bmin, bmax, Q)
    if order == 1:
        R, J = cell_residual(p, pos, mass)
    else:
        R, J = cell_remainder(p, pos, mass)
        scale = scale * f[0]
    Jv = np.array([J[0, 0], J[1, 1], J[2, 2], J[0, 1], J[0, 2], J[1, 2]])
    return f, b, np.concatenate([R / scale, K_JAC * Jv * d / scale])


def quadrupole_coefficients(f):
    """Basis coefficients of the exact quadrupole term (features f)."""
    trq, uqu = f[..., 2], f[..., 3]
    c = np.zeros(np.shape(f)[:-1] + (K_OUT,))
    c[..., 0] = 7.5 * uqu - 1.5 * trq
    c[..., 1] = -3.0
    c[..., 3] = 3.0 * K_JAC
    c[..., 4] = -15.0 * K_JAC
    c[..., 5] = (1.5 * trq - 7.5 * uqu) * K_JAC
    c[..., 6] = (52.5 * uqu - 7.5 * trq) * K_JAC
    return c


def quadrupole_field(p, cm, Q):
    """Textbook quadrupole term of the field of second moments Q about cm at p:
    a = -grad Phi2, Phi2 = -(3 n.Q.n - tr Q) / (2 d^3), n = (p - cm)/d."""
    s = np.asarray(p, np.float64) - np.asarray(cm, np.float64)
    d = np.linalg.norm(s)
    n = s / d
    Qm = np.array([[Q[0], Q[3], Q[4]], [Q[3], Q[1], Q[5]], [Q[4], Q[5], Q[2]]])
    qn = Qm @ n
    nqn = n @ qn
    tr = np.trace(Qm)
    # grad of -(3 nQn - tr)/(2 d^3) with respect to p
    grad = -(3.0 * (2.0 * qn - 2.0 * nqn * n) / d) / (2 * d ** 3) + 3.0 * (3 * nqn - tr) * n / (2 * d ** 4)
    return -grad


def group_band(n, group, cm, crit, bmin, bmax, rho_min, near_k=0.0):
    """(far, taken, near) heap nodes of group `group` of a heap tree of n bodies.

    cm[2n][3], crit[2n], bmin/bmax[2n][3] as returned by nbh_tree_get.  A node is
    taken when the distance from the group box to its mass centre squared exceeds
    crit (leaves always).  A taken internal node whose mass centre lies within
    near_k half-diagonals of the box centre is near; otherwise it is far when its
    bounding radius exceeds rho_min times that distance (thresholds rounded to
    float32 as the kernels receive them)."""
    g = n // 64 + group
    lo, hi = bmin[g], bmax[g]
    c = 0.5 * (lo + hi)
    half = 0.5 * np.linalg.norm(hi - lo)
    rho = float(np.float32(rho_min))
    nk = float(np.float32(near_k))
    far, taken, near = [], [], []
    stack = [1]
    while stack:
        node = stack.pop()
        if node >= n:
            taken.append(node)
            continue
        e = np.maximum(np.maximum(lo - cm[node], cm[node] - hi), 0.0)
        if e @ e > crit[node]:
            taken.append(node)
            u = cm[node] - c
            dc = np.sqrt(u @ u)
            if dc < nk * half:
                near.append(node)
            elif cell_radius(bmin[node], bmax[node], cm[node]) > rho * dc:
                far.append(node)
        else:
            stack += [2 * node + 1, 2 * node]
    return far, taken, near


def quadrupole_fields(P, cm, Q):
    """Sum over cells of quadrupole_field(P[b], cm[c], Q[c]) for points P[b,3]."""
    P = np.atleast_2d(np.asarray(P, np.float64))
    cm = np.atleast_2d(np.asarray(cm, np.float64))
    Q = np.atleast_2d(np.asarray(Q, np.float64))
    if cm.shape[0] == 0:
        return np.zeros_like(P)
    Qm = np.zeros((Q.shape[0], 3, 3))
    for (i, j), k in zip(((0, 0), (1, 1), (2, 2), (0, 1), (0, 2), (1, 2)), range(6)):
        Qm[:, i, j] = Q[:, k]
        Qm[:, j, i] = Q[:, k]
    s = P[:, None, :] - cm[None, :, :]                      # [b,c,3], from the cell to the point
    d = np.linalg.norm(s, axis=2)
    nv = s / d[..., None]
    qn = np.einsum("cij,bcj->bci", Qm, nv)
    nqn = np.einsum("bci,bci->bc", nv, qn)
    tr = np.trace(Qm, axis1=1, axis2=2)[None, :]
    grad = (-(3.0 * (2.0 * qn - 2.0 * nqn[..., None] * nv) / d[..., None]) / (2 * d[..., None] ** 3)
            + 3.0 * (3 * nqn - tr)[..., None] * nv / (2 * d[..., None] ** 4))
    return -grad.sum(1)


def local_expansion_tensors(R, M, order: int = 2):
    """Taylor coefficients, about an expansion centre, of the monopole fields of
    cells of mass M[c] whose mass centres sit at R[c,3] from the centre.

    With f(R) = M R / d^3 the field at centre + delta is f(R - delta), so
      a(delta) = f - D1 delta + 1/2 D2[delta, delta] - 1/6 D3[delta, delta, delta] + ...,
      D1_ij = M (delta_ij / d^3 - 3 R_i R_j / d^5),
      D2_ijk = M (-3 (delta_ij R_k + delta_ik R_j + delta_jk R_i) / d^5 + 15 R_i R_j R_k / d^7),
      D3_ijkl = M (-3 (delta_ij delta_kl + delta_ik delta_jl + delta_il delta_jk) / d^5
                   + 15 (delta_ij R_k R_l + delta_ik R_j R_l + delta_il R_j R_k
                         + delta_jk R_i R_l + delta_jl R_i R_k + delta_kl R_i R_j) / d^7
                   - 105 R_i R_j R_k R_l / d^9).
    Returns the sums (f[3], J = -D1 [3,3], D2 [3,3,3]) and, for order 3, D3 [3,3,3,3]."""
    R = np.atleast_2d(np.asarray(R, np.float64))
    M = np.atleast_1d(np.asarray(M, np.float64))
    d = np.linalg.norm(R, axis=1)
    I = np.eye(3)
    w3, w5, w7, w9 = M / d ** 3, M / d ** 5, M / d ** 7, M / d ** 9
    f = np.einsum("c,ci->i", w3, R)
    D1 = w3.sum() * I - 3.0 * np.einsum("c,ci,cj->ij", w5, R, R)
    D2 = (-3.0 * (np.einsum("ij,k->ijk", I, w5 @ R) + np.einsum("ik,j->ijk", I, w5 @ R)
                  + np.einsum("jk,i->ijk", I, w5 @ R))
          + 15.0 * np.einsum("c,ci,cj,ck->ijk", w7, R, R, R))
    if order < 3:
        return f, -D1, D2
    A = np.einsum("c,ci,cj->ij", w7, R, R)
    D3 = (-3.0 * w5.sum() * (np.einsum("ij,kl->ijkl", I, I) + np.einsum("ik,jl->ijkl", I, I)
                             + np.einsum("il,jk->ijkl", I, I))
          + 15.0 * (np.einsum("ij,kl->ijkl", I, A) + np.einsum("ik,jl->ijkl", I, A) + np.einsum("il,jk->ijkl", I, A)
                    + np.einsum("jk,il->ijkl", I, A) + np.einsum("jl,ik->ijkl", I, A)
                    + np.einsum("kl,ij->ijkl", I, A))
          - 105.0 * np.einsum("c,ci,cj,ck,cl->ijkl", w9, R, R, R, R))
    return f, -D1, D2, D3


def local_expansion_eval(f, J, D2, delta, D3=None):
    """a(delta) = f + J delta + 1/2 D2[delta, delta] (- 1/6 D3[delta, delta, delta])
    for offsets delta[b,3]."""
    delta = np.atleast_2d(np.asarray(delta, np.float64))
    a = f[None, :] + delta @ J.T + 0.5 * np.einsum("ijk,bj,bk->bi", D2, delta, delta)
    if D3 is not None:
        a -= np.einsum("ijkl,bj,bk,bl->bi", D3, delta, delta, delta) / 6.0
    return a


def _box(bmin, bmax, node):
    lo, hi = bmin[node], bmax[node]
    return lo, hi, 0.5 * (lo + hi), 0.5 * np.linalg.norm(hi - lo)


def super_band(n, s, cm, crit, bmin, bmax, near_k, ratio, shift=4):
    """(far, listed) heap nodes of super-group s (heap node n / 64 / 2^shift + s)
    of the two-level local walk: walking from the root with the super-group's
    box, an accepted internal node whose mass centre lies at least near_k
    half-diagonals from the box centre is far (expanded about that centre), any
    other accepted node is listed; a rejected node is listed when the ball about
    its mass centre that holds all its bodies (radius sqrt(crit) / ratio, the
    cell radius of the opening rule) lies within that distance of the centre,
    and opened otherwise."""
    sg = (n >> (6 + shift)) + s
    lo, hi, c, half = _box(bmin, bmax, sg)
    nk = float(np.float32(near_k))
    far, listed = [], []
    stack = [1]
    while stack:
        node = stack.pop()
        if node >= n:
            listed.append(node)
            continue
        u = cm[node] - c
        du = np.sqrt(u @ u)
        e = np.maximum(np.maximum(lo - cm[node], cm[node] - hi), 0.0)
        if e @ e > crit[node]:
            (listed if du < nk * half else far).append(node)
        elif du + np.sqrt(crit[node]) / ratio < nk * half:
            listed.append(node)
        else:
            stack += [2 * node + 1, 2 * node]
    return far, listed


def group_band_from(n, group, start, cm, crit, bmin, bmax, near_k):
    """group_band (rho_min 0) for a walk that starts from the nodes `start`
    instead of the root: (far, taken, near)."""
    g = n // 64 + group
    lo, hi, c, half = _box(bmin, bmax, g)
    nk = float(np.float32(near_k))
    far, taken, near = [], [], []
    stack = list(start)[::-1]
    while stack:
        node = stack.pop()
        if node >= n:
            taken.append(node)
            continue
        e = np.maximum(np.maximum(lo - cm[node], cm[node] - hi), 0.0)
        if e @ e > crit[node]:
            taken.append(node)
            u = cm[node] - c
            (near if np.sqrt(u @ u) < nk * half else far).append(node)
        else:
            stack += [2 * node + 1, 2 * node]
    return far, taken, near
