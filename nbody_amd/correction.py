"""Training and storage of the learned Barnes-Hut correction (no reference
counterpart).

``train`` builds the heap tree of a body state on the GPU, collects the far
cells of a random subset of 64-body groups with their exact residual fields
(nbh_bh_band_samples), and fits the network with the device's own forward,
backward and Adam kernels (nbh_mlp_train).  ``train_two_level`` does the same
for the two-level local walk: the far cells of the super-groups with their
exact field remainder after the monopole and quadrupole terms
(nbh_bh_super_samples).  ``save`` writes the file read by the ``hip_bh``
engine's ``learned_correction`` key:

    8 bytes of magic -- "NBHMLP01" (group walk) or "NBHMLP2L" (two-level
    walk) --, int32 parameter count, float32 rho_min, float32 near_k, then the
    float32 parameters (all little-endian).
"""
from __future__ import annotations

import ctypes as C
import struct
from dataclasses import dataclass, field
from pathlib import Path

import numpy as np

from . import native

MAGIC = b"NBHMLP01"
MAGIC_TWO_LEVEL = b"NBHMLP2L"


@dataclass
class Correction:
    params: np.ndarray            # float32 [param_count]
    rho_min: float = 0.0
    near_k: float = 2.0
    stats: dict = field(default_factory=dict)
    kind: str = "group"           # "group" or "two_level"


def param_count() -> int:
    c = C.c_int()
    native.check(native.hip_lib().nbh_mlp_param_count(C.byref(c)), "nbh_mlp_param_count")
    return c.value


def save(path, corr: Correction) -> None:
    p = np.ascontiguousarray(corr.params, dtype="<f4")
    if p.size != param_count():
        raise ValueError(f"expected {param_count()} parameters, got {p.size}")
    with open(path, "wb") as fh:
        fh.write(MAGIC_TWO_LEVEL if corr.kind == "two_level" else MAGIC)
        fh.write(struct.pack("<iff", p.size, float(corr.rho_min), float(corr.near_k)))
        fh.write(p.tobytes())


def load(path) -> Correction:
    raw = Path(path).read_bytes()
    if raw[:8] not in (MAGIC, MAGIC_TWO_LEVEL) or len(raw) < 20:
        raise ValueError(f"{path}: not a correction network file")
    count, rho, near = struct.unpack("<iff", raw[8:20])
    if count != param_count() or len(raw) != 20 + 4 * count:
        raise ValueError(f"{path}: wrong parameter count")
    return Correction(np.frombuffer(raw[20:], dtype="<f4").astype(np.float32), rho, near,
                      kind="two_level" if raw[:8] == MAGIC_TWO_LEVEL else "group")


def train(y: np.ndarray, mass: np.ndarray, ratio: float, rho_min: float = 0.0, near_k: float = 2.0,
          max_samples: int = 2_000_000, epochs: int = 30, batch: int = 1024, lr: float = 2e-3,
          seed: int = 0, device: int = 0) -> Correction:
    """Fit the correction on the state y[6n] (SoA), mass[n]; n a power of two >= 64.

    Groups are drawn at random until about max_samples far cells are collected."""
    lib = native.hip_lib()
    n = mass.size
    if n < 64 or n & (n - 1):
        raise ValueError("n must be a power of two >= 64")
    ngroups = n // 64
    ctx = C.c_void_p()
    native.check(lib.nbh_ctx_create(device, C.byref(ctx)), "nbh_ctx_create")
    bufs, tree, net = [], C.c_void_p(), C.c_void_p()

    def dev(nbytes):
        p = C.c_void_p()
        native.check(lib.nbh_malloc(ctx, max(8, nbytes), C.byref(p)), "nbh_malloc")
        bufs.append(p)
        return p

    try:
        y = np.ascontiguousarray(y, dtype=np.float64)
        m = np.ascontiguousarray(mass, dtype=np.float64)
        dy, dm = dev(y.nbytes), dev(m.nbytes)
        native.check(lib.nbh_memcpy_h2d(ctx, dy, y.ctypes.data_as(C.c_void_p), y.nbytes), "h2d")
        native.check(lib.nbh_memcpy_h2d(ctx, dm, m.ctypes.data_as(C.c_void_p), m.nbytes), "h2d")
        native.check(lib.nbh_tree_create(ctx, n, C.byref(tree)), "nbh_tree_create")
        native.check(lib.nbh_tree_build(ctx, tree, dy, dm, ratio), "nbh_tree_build")
        rng = np.random.default_rng(seed)
        order = rng.permutation(ngroups).astype(np.uint32)
        # far cells per group from a small probe, then enough groups for max_samples
        probe = order[:min(ngroups, 16)]
        cnt = C.c_size_t()
        native.check(lib.nbh_bh_band_samples(ctx, tree, probe.ctypes.data_as(C.c_void_p), probe.size, rho_min,
                                             near_k, None, None, None, None, 0, C.byref(cnt)), "band samples")
        per_group = max(1.0, cnt.value / probe.size)
        take = int(min(ngroups, max(1, round(max_samples / per_group))))
        groups = np.ascontiguousarray(order[:take])
        cap = int(max_samples * 1.25) + 1024
        dx, db, dt = dev(4 * 16 * cap), dev(4 * 16 * cap), dev(4 * 9 * cap)
        native.check(lib.nbh_bh_band_samples(ctx, tree, groups.ctypes.data_as(C.c_void_p), groups.size, rho_min,
                                             near_k, dx, db, dt, None, cap, C.byref(cnt)), "band samples")
        count = min(cnt.value, cap)
        if count == 0:
            raise RuntimeError("no far cells: nothing to train on")
        native.check(lib.nbh_mlp_create(ctx, seed, C.byref(net)), "nbh_mlp_create")
        el = np.empty(epochs, dtype=np.float32)
        native.check(lib.nbh_mlp_train(ctx, net, dx, db, dt, count, epochs, batch, lr,
                                       el.ctypes.data_as(C.c_void_p)), "nbh_mlp_train")
        params = np.empty(param_count(), dtype=np.float32)
        native.check(lib.nbh_mlp_get_params(ctx, net, params.ctypes.data_as(C.c_void_p)), "get params")
        return Correction(params, rho_min, near_k,
                          {"samples": int(count), "groups": int(take), "epochs": epochs,
                           "loss_first": float(el[0]), "loss_last": float(el[-1])})
    finally:
        if net:
            lib.nbh_mlp_destroy(net)
        if tree:
            lib.nbh_tree_destroy(tree)
        for p in bufs:
            lib.nbh_free(ctx, p)
        lib.nbh_ctx_destroy(ctx)


def train_two_level(y: np.ndarray, mass: np.ndarray, ratio: float, near_k: float = 5.0,
                    max_samples: int = 2_000_000, epochs: int = 40, batch: int = 1024, lr: float = 2e-3,
                    seed: int = 0, device: int = 0) -> Correction:
    """Fit the two-level walk's correction on the state y[6n] (SoA), mass[n];
    n a power of two >= 1024.  Super-groups are drawn at random until about
    max_samples far cells are collected."""
    lib = native.hip_lib()
    n = mass.size
    if n < 1024 or n & (n - 1):
        raise ValueError("n must be a power of two >= 1024")
    nsup = n >> 10
    ctx = C.c_void_p()
    native.check(lib.nbh_ctx_create(device, C.byref(ctx)), "nbh_ctx_create")
    bufs, tree, net = [], C.c_void_p(), C.c_void_p()

    def dev(nbytes):
        p = C.c_void_p()
        native.check(lib.nbh_malloc(ctx, max(8, nbytes), C.byref(p)), "nbh_malloc")
        bufs.append(p)
        return p

    try:
        y = np.ascontiguousarray(y, dtype=np.float64)
        m = np.ascontiguousarray(mass, dtype=np.float64)
        dy, dm = dev(y.nbytes), dev(m.nbytes)
        native.check(lib.nbh_memcpy_h2d(ctx, dy, y.ctypes.data_as(C.c_void_p), y.nbytes), "h2d")
        native.check(lib.nbh_memcpy_h2d(ctx, dm, m.ctypes.data_as(C.c_void_p), m.nbytes), "h2d")
        native.check(lib.nbh_tree_create(ctx, n, C.byref(tree)), "nbh_tree_create")
        native.check(lib.nbh_tree_build(ctx, tree, dy, dm, ratio), "nbh_tree_build")
        rng = np.random.default_rng(seed)
        order = rng.permutation(nsup).astype(np.uint32)
        probe = np.ascontiguousarray(order[:min(nsup, 16)])
        cnt = C.c_size_t()
        native.check(lib.nbh_bh_super_samples(ctx, tree, near_k, probe.ctypes.data_as(C.c_void_p), probe.size,
                                              None, None, None, None, 0, C.byref(cnt)), "super samples")
        per_super = max(1.0, cnt.value / probe.size)
        take = int(min(nsup, max(1, round(max_samples / per_super))))
        sel = np.ascontiguousarray(order[:take])
        cap = int(max_samples * 1.25) + 1024
        dx, db, dt = dev(4 * 16 * cap), dev(4 * 16 * cap), dev(4 * 9 * cap)
        native.check(lib.nbh_bh_super_samples(ctx, tree, near_k, sel.ctypes.data_as(C.c_void_p), sel.size,
                                              dx, db, dt, None, cap, C.byref(cnt)), "super samples")
        count = min(cnt.value, cap)
        if count == 0:
            raise RuntimeError("no super-group far cells: nothing to train on")
        native.check(lib.nbh_mlp_create(ctx, seed, C.byref(net)), "nbh_mlp_create")
        el = np.empty(epochs, dtype=np.float32)
        native.check(lib.nbh_mlp_train(ctx, net, dx, db, dt, count, epochs, batch, lr,
                                       el.ctypes.data_as(C.c_void_p)), "nbh_mlp_train")
        params = np.empty(param_count(), dtype=np.float32)
        native.check(lib.nbh_mlp_get_params(ctx, net, params.ctypes.data_as(C.c_void_p)), "get params")
        return Correction(params, 0.0, near_k,
                          {"samples": int(count), "super_groups": int(take), "epochs": epochs,
                           "loss_first": float(el[0]), "loss_last": float(el[-1])}, kind="two_level")
    finally:
        if net:
            lib.nbh_mlp_destroy(net)
        if tree:
            lib.nbh_tree_destroy(tree)
        for p in bufs:
            lib.nbh_free(ctx, p)
        lib.nbh_ctx_destroy(ctx)
