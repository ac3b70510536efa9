"""ctypes access to the CPU oracle (oracle/liboracle.so, oracle/liboracle_engine.so).

Test infrastructure: the oracle is the checker the HIP engines are compared
against; nothing in the product imports this module.
"""
from __future__ import annotations

import ctypes as C
from pathlib import Path

import numpy as np

import nbody_amd
from nbody_amd.api import nbody_engine

ROOT = Path(__file__).resolve().parent.parent
LIB = ROOT / "oracle" / "liboracle.so"
LIB_ENGINE = ROOT / "oracle" / "liboracle_engine.so"

PD = C.POINTER(C.c_double)
sz = C.c_size_t
_lib = None
_elib = None


class ah_t(C.Structure):
    _fields_ = [("n", sz), ("max_dist_sqr", C.c_double), ("min_force_sqr", C.c_double),
                ("offs", C.POINTER(sz)), ("list", C.POINTER(C.c_uint)), ("univ", PD), ("total", sz)]


class heap_t(C.Structure):
    _fields_ = [("count", sz), ("size", sz), ("cm", PD), ("mass", PD), ("crit_r2", PD),
                ("bmin", PD), ("bmax", PD), ("body_n", C.POINTER(sz)), ("ratio_sqr", C.c_double)]


def lib():
    global _lib
    if _lib is None:
        if not LIB.exists():
            raise RuntimeError("oracle not built: run `python -m nbody_amd.build`")
        L = C.CDLL(str(LIB))
        sig = {
            "oracle_force": (None, [PD, PD, C.c_double, C.c_double, PD]),
            "oracle_potential_energy": (C.c_double, [PD, PD, C.c_double, C.c_double]),
            "oracle_fcompute_simple": (None, [sz, PD, PD, PD]),
            "oracle_fcompute_openmp": (None, [sz, PD, PD, PD, C.c_int]),
            "oracle_accel_targets": (None, [sz, PD, PD, C.POINTER(sz), sz, PD, C.c_int]),
            "oracle_fcompute_block": (None, [sz, PD, PD, PD, C.c_int]),
            "oracle_ah_alloc": (C.c_int, [C.POINTER(ah_t), sz, C.c_double, C.c_double]),
            "oracle_ah_free": (None, [C.POINTER(ah_t)]),
            "oracle_ah_full": (C.c_int, [C.POINTER(ah_t), PD, PD, PD]),
            "oracle_ah_sparse": (None, [C.POINTER(ah_t), PD, PD, PD]),
            "oracle_fill": (None, [PD, C.c_double, sz]),
            "oracle_fmadd_inplace": (None, [PD, PD, C.c_double, sz]),
            "oracle_fmadd": (None, [PD, PD, PD, C.c_double, sz]),
            "oracle_fmaddn_inplace": (None, [PD, C.POINTER(PD), PD, sz, sz]),
            "oracle_fmaddn": (None, [PD, PD, C.POINTER(PD), PD, sz, sz]),
            "oracle_fmaddn_corr": (None, [PD, PD, C.POINTER(PD), PD, sz, sz]),
            "oracle_fmaxabs": (C.c_double, [PD, sz]),
            "oracle_clamp": (None, [PD, sz, C.c_double]),
            "oracle_heap_left": (sz, [sz]),
            "oracle_heap_right": (sz, [sz]),
            "oracle_heap_parent": (sz, [sz]),
            "oracle_heap_next_down": (sz, [sz]),
            "oracle_heap_skip_idx": (sz, [sz]),
            "oracle_heap_next_up": (sz, [sz, sz]),
            "oracle_heap_alloc": (C.c_int, [C.POINTER(heap_t), sz]),
            "oracle_heap_free": (None, [C.POINTER(heap_t)]),
            "oracle_heap_build": (C.c_int, [C.POINTER(heap_t), PD, PD, PD, PD, C.c_double]),
            "oracle_heap_rebuild": (None, [C.POINTER(heap_t), PD, PD, PD, C.c_double]),
            "oracle_heap_build_omp": (C.c_int, [C.POINTER(heap_t), PD, PD, PD, PD, C.c_double, C.c_int]),
            "oracle_bh_accel_targets": (None, [C.POINTER(heap_t), C.c_int, sz, PD, PD, C.POINTER(sz), sz, PD,
                                               C.c_int]),
            "oracle_heap_traverse": (None, [C.POINTER(heap_t), PD, C.c_double, PD]),
            "oracle_heap_traverse_stackless": (None, [C.POINTER(heap_t), PD, C.c_double, PD]),
            "oracle_bh_fcompute": (None, [C.POINTER(heap_t), C.c_int, C.c_int, sz, PD, PD, PD, C.c_int]),
            "oracle_stats": (None, [sz, PD, PD, PD, C.c_int, PD]),
            "oracle_radix_build": (C.c_int, [sz, PD, PD, PD, PD, C.c_double, C.c_void_p, C.c_void_p, C.c_void_p,
                                             PD, PD, PD, PD, PD, PD]),
        }
        for k, (r, a) in sig.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def elib():
    global _elib
    if _elib is None:
        nbody_amd.host_lib()
        if not LIB_ENGINE.exists():
            raise RuntimeError("oracle engine not built: run `python -m nbody_amd.build`")
        L = C.CDLL(str(LIB_ENGINE))
        L.oracle_engine_create_simple.restype = C.c_void_p
        L.oracle_engine_create_simple.argtypes = [C.c_int]
        L.oracle_engine_create_simple_bh.restype = C.c_void_p
        L.oracle_engine_create_simple_bh.argtypes = [C.c_double, C.c_int, C.c_int, sz, C.c_int]
        L.oracle_engine_create_block.restype = C.c_void_p
        L.oracle_engine_create_block.argtypes = [C.c_int]
        L.oracle_engine_create_ah.restype = C.c_void_p
        L.oracle_engine_create_ah.argtypes = [sz, C.c_double, C.c_double]
        L.oracle_engine_ah_lists.restype = sz
        L.oracle_engine_ah_lists.argtypes = [C.c_void_p, C.POINTER(sz), C.POINTER(C.c_uint)]
        L.oracle_engine_heap.restype = C.c_int
        L.oracle_engine_heap.argtypes = [C.c_void_p, C.POINTER(C.c_int64), PD, PD, PD, PD, PD]
        _elib = L
    return _elib


def p(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(PD)


def simple_engine(threads: int = 1) -> nbody_engine:
    """The reference's `simple` engine (threads=1) or `openmp` engine (threads>1)."""
    return nbody_engine(elib().oracle_engine_create_simple(threads))


def block_engine(threads: int = 1) -> nbody_engine:
    """The reference's `block` engine (nbody/nbody_engine_block.cpp:14-126)."""
    return nbody_engine(elib().oracle_engine_create_block(threads))


def ah_engine(rate: int = 1000, max_dist: float = 10.0, min_force: float = 1e-4) -> nbody_engine:
    """The reference's `ah` engine (nbody/nbody_engine_ah.cpp:3-160), same defaults."""
    return nbody_engine(elib().oracle_engine_create_ah(rate, max_dist, min_force))


def ah_engine_lists(e: nbody_engine, n: int):
    """(offsets[n + 1], list) of the last full recompute of an oracle `ah` engine."""
    total = elib().oracle_engine_ah_lists(e.handle, None, None)
    offs = np.empty(n + 1, dtype=np.uint64)
    lst = np.empty(max(total, 1), dtype=np.uint32)
    elib().oracle_engine_ah_lists(e.handle, offs.ctypes.data_as(C.POINTER(sz)), lst.ctypes.data_as(C.POINTER(C.c_uint)))
    return offs, lst[:total]


def simple_bh_engine(ratio: float, layout: str = "heap", traverse: str = "nested_tree",
                     rate: int = 0, threads: int = 0) -> nbody_engine:
    lay = {"heap": 1, "heap_stackless": 2}[layout]
    tt = {"cycle": 0, "nested_tree": 1}[traverse]
    return nbody_engine(elib().oracle_engine_create_simple_bh(ratio, lay, tt, rate, threads))


def engine_heap(e: nbody_engine, n: int):
    size = 2 * n
    body = np.empty(size, dtype=np.int64)
    cm = np.empty(size * 3); mass = np.empty(size); crit = np.empty(size)
    bmin = np.empty(size * 3); bmax = np.empty(size * 3)
    ok = elib().oracle_engine_heap(e.handle, body.ctypes.data_as(C.POINTER(C.c_int64)), p(cm), p(mass),
                                   p(crit), p(bmin), p(bmax))
    assert ok
    return dict(body_n=body, cm=cm.reshape(-1, 3), mass=mass, crit_r2=crit,
                bmin=bmin.reshape(-1, 3), bmax=bmax.reshape(-1, 3))


def fcompute_simple(y: np.ndarray, mass: np.ndarray, threads: int = 1) -> np.ndarray:
    n = len(mass)
    f = np.empty(6 * n)
    if threads == 1:
        lib().oracle_fcompute_simple(n, p(y), p(mass), p(f))
    else:
        lib().oracle_fcompute_openmp(n, p(y), p(mass), p(f), threads)
    return f


def fcompute_block(y: np.ndarray, mass: np.ndarray, threads: int = 1) -> np.ndarray:
    n = len(mass)
    f = np.empty(6 * n)
    lib().oracle_fcompute_block(n, p(y), p(mass), p(f), threads)
    return f


class AH:
    """oracle_ah owner: the Ahmad-Cohen full and sparse passes."""

    def __init__(self, n: int, max_dist: float, min_force: float):
        self.h = ah_t()
        assert lib().oracle_ah_alloc(C.byref(self.h), n, max_dist, min_force) == 0
        self.n = n

    def __del__(self):
        lib().oracle_ah_free(C.byref(self.h))

    def full(self, y: np.ndarray, mass: np.ndarray) -> np.ndarray:
        f = np.empty(6 * self.n)
        assert lib().oracle_ah_full(C.byref(self.h), p(y), p(np.ascontiguousarray(mass)), p(f)) == 0
        return f

    def sparse(self, y: np.ndarray, mass: np.ndarray) -> np.ndarray:
        f = np.empty(6 * self.n)
        lib().oracle_ah_sparse(C.byref(self.h), p(y), p(np.ascontiguousarray(mass)), p(f))
        return f

    def lists(self):
        n, total = self.n, self.h.total
        offs = np.ctypeslib.as_array(self.h.offs, shape=(n + 1,)).astype(np.uint64)
        lst = np.ctypeslib.as_array(self.h.list, shape=(total,)).copy() if total else np.empty(0, np.uint32)
        univ = np.ctypeslib.as_array(self.h.univ, shape=(3 * n,)).reshape(n, 3).copy()
        return offs, lst, univ


def accel_targets(y: np.ndarray, mass: np.ndarray, idx: np.ndarray, threads: int = 0) -> np.ndarray:
    idx = np.ascontiguousarray(idx, dtype=np.uint64)
    out = np.empty(3 * len(idx))
    lib().oracle_accel_targets(len(mass), p(y), p(mass), idx.ctypes.data_as(C.POINTER(sz)), len(idx),
                               p(out), threads)
    return out.reshape(-1, 3)


class Heap:
    """oracle_heap owner."""

    def __init__(self, n: int):
        self.h = heap_t()
        assert lib().oracle_heap_alloc(C.byref(self.h), n) == 0
        self.n = n

    def __del__(self):
        lib().oracle_heap_free(C.byref(self.h))

    def build(self, y: np.ndarray, mass: np.ndarray, ratio: float):
        n = self.n
        rx = np.ascontiguousarray(y[:n]); ry = np.ascontiguousarray(y[n:2 * n]); rz = np.ascontiguousarray(y[2 * n:3 * n])
        lib().oracle_heap_build(C.byref(self.h), p(rx), p(ry), p(rz), p(np.ascontiguousarray(mass)), ratio)

    def build_omp(self, y: np.ndarray, mass: np.ndarray, ratio: float, threads: int = 0):
        n = self.n
        rx = np.ascontiguousarray(y[:n]); ry = np.ascontiguousarray(y[n:2 * n]); rz = np.ascontiguousarray(y[2 * n:3 * n])
        assert lib().oracle_heap_build_omp(C.byref(self.h), p(rx), p(ry), p(rz), p(np.ascontiguousarray(mass)),
                                           ratio, threads) == 0

    def accel_targets(self, y, mass, idx, layout: int = 2, threads: int = 0) -> np.ndarray:
        idx = np.ascontiguousarray(idx, dtype=np.uint64)
        out = np.empty(3 * len(idx))
        lib().oracle_bh_accel_targets(C.byref(self.h), layout, self.n, p(np.ascontiguousarray(y)),
                                      p(np.ascontiguousarray(mass)), idx.ctypes.data_as(C.POINTER(sz)), len(idx),
                                      p(out), threads)
        return out.reshape(-1, 3)

    def rebuild(self, y: np.ndarray, ratio: float):
        n = self.n
        rx = np.ascontiguousarray(y[:n]); ry = np.ascontiguousarray(y[n:2 * n]); rz = np.ascontiguousarray(y[2 * n:3 * n])
        lib().oracle_heap_rebuild(C.byref(self.h), p(rx), p(ry), p(rz), ratio)

    def arrays(self):
        s = self.h.size
        # size_t(-1) marks internal nodes; viewed as int64 it reads -1
        body = np.ctypeslib.as_array(self.h.body_n, shape=(s,)).view(np.int64).copy()
        return dict(body_n=body,
                    cm=np.ctypeslib.as_array(self.h.cm, shape=(s * 3,)).reshape(-1, 3).copy(),
                    mass=np.ctypeslib.as_array(self.h.mass, shape=(s,)).copy(),
                    crit_r2=np.ctypeslib.as_array(self.h.crit_r2, shape=(s,)).copy(),
                    bmin=np.ctypeslib.as_array(self.h.bmin, shape=(s * 3,)).reshape(-1, 3).copy(),
                    bmax=np.ctypeslib.as_array(self.h.bmax, shape=(s * 3,)).reshape(-1, 3).copy())

    def traverse(self, v1, mass1: float, stackless: bool = False) -> np.ndarray:
        v = np.ascontiguousarray(v1, dtype=np.float64)
        out = np.empty(3)
        fn = lib().oracle_heap_traverse_stackless if stackless else lib().oracle_heap_traverse
        fn(C.byref(self.h), p(v), mass1, p(out))
        return out

    def fcompute(self, y, mass, layout: int = 2, traverse: int = 1, threads: int = 0) -> np.ndarray:
        f = np.empty(6 * self.n)
        lib().oracle_bh_fcompute(C.byref(self.h), layout, traverse, self.n, p(y), p(np.ascontiguousarray(mass)),
                                 p(f), threads)
        return f


def radix_build(y: np.ndarray, mass: np.ndarray, ratio: float) -> dict:
    """The oracle's Morton radix tree of the state y[6n] (SoA): arrays by node id
    (root 1, internal 1..n-1, sorted leaf j -> n + j)."""
    n = len(mass)
    rx, ry, rz = (np.ascontiguousarray(y[k * n:(k + 1) * n]) for k in range(3))
    out = dict(order=np.empty(n, np.uint32), child=np.empty(2 * n, np.int32), parent=np.empty(2 * n, np.int32),
               cm=np.zeros(6 * n), mass=np.zeros(2 * n), crit=np.zeros(2 * n), bmin=np.zeros(6 * n),
               bmax=np.zeros(6 * n), quad=np.zeros(12 * n))
    rc = lib().oracle_radix_build(n, p(rx), p(ry), p(rz), p(np.ascontiguousarray(mass)), ratio,
                                  out["order"].ctypes.data_as(C.c_void_p), out["child"].ctypes.data_as(C.c_void_p),
                                  out["parent"].ctypes.data_as(C.c_void_p), p(out["cm"]), p(out["mass"]),
                                  p(out["crit"]), p(out["bmin"]), p(out["bmax"]), p(out["quad"]))
    assert rc == 0
    for k in ("cm", "bmin", "bmax"):
        out[k] = out[k].reshape(-1, 3)
    out["quad"] = out["quad"].reshape(-1, 6)
    out["child"] = out["child"].reshape(-1, 2)
    return out


def stats(pos: np.ndarray, vel: np.ndarray, mass: np.ndarray, energy: bool = True) -> np.ndarray:
    out = np.empty(12)
    lib().oracle_stats(len(mass), p(np.ascontiguousarray(pos)), p(np.ascontiguousarray(vel)),
                       p(np.ascontiguousarray(mass)), 1 if energy else 0, p(out))
    return out


def state_from_data(d) -> tuple[np.ndarray, np.ndarray]:
    """(y[6n] SoA, mass[n]) of an nbody_data."""
    pos, vel, m, _ = d.bodies()
    y = np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2]])
    return np.ascontiguousarray(y), np.ascontiguousarray(m)
