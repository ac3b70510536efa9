"""Python mirror of the reference's engine/solver/data interface.

Same names and argument meaning as nbody_data, nbody_engine, nbody_solver and
the factories nbody_create_engine / nbody_create_solver of drons/nbody
(nbody/nbody_engine.h:16-96, nbody/nbody_engines.cpp:4-193,
nbody/nbody_solvers.cpp:17-110).  Every call goes to the C++ host layer in
libnbody.so; buffers are numpy float64 arrays on the host side.
"""
from __future__ import annotations

import ctypes as C
from typing import Optional, Sequence

import numpy as np

from . import native

_PD = C.POINTER(C.c_double)


def _pd(a: Optional[np.ndarray]):
    if a is None:
        return None
    return a.ctypes.data_as(_PD)


def _params(d: dict) -> bytes:
    parts = []
    for k, v in d.items():
        if isinstance(v, bool):
            v = "true" if v else "false"
        parts.append(f"{k}={v}")
    return ";".join(parts).encode()


def log_count() -> int:
    return native.host_lib().nbs_log_count()


def log_last() -> str:
    return native.host_lib().nbs_log_last().decode(errors="replace")


def log_quiet(q: bool = True) -> None:
    native.host_lib().nbs_log_quiet(1 if q else 0)


class nbody_data:
    """Bodies, time, step and the P/L/E/V checks (nbody/nbody_data.h)."""

    def __init__(self):
        self._lib = native.host_lib()
        self.handle = self._lib.nbs_data_create()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self._lib.nbs_data_destroy(h)
            self.handle = None

    def load(self, path: str, kind: str = "G1") -> bool:
        return bool(self._lib.nbs_data_load(self.handle, str(path).encode(), kind.encode()))

    def load_initial(self, path: str, kind: str) -> bool:
        return self.load(path, kind)

    def save(self, path: str) -> bool:
        return bool(self._lib.nbs_data_save(self.handle, str(path).encode()))

    def make_universe(self, star_count: int, sx: float, sy: float, sz: float) -> None:
        self._lib.nbs_data_make_universe(self.handle, star_count, sx, sy, sz)

    def generate(self, kind: str, n: int, seed: int) -> None:
        if not self._lib.nbs_data_generate(self.handle, kind.encode(), n, seed):
            raise ValueError(f"unknown generator {kind!r}")

    def set_bodies(self, pos: np.ndarray, vel: np.ndarray, mass: np.ndarray,
                   radius: Optional[np.ndarray] = None) -> None:
        pos = np.ascontiguousarray(pos, dtype=np.float64).reshape(-1, 3)
        vel = np.ascontiguousarray(vel, dtype=np.float64).reshape(-1, 3)
        mass = np.ascontiguousarray(mass, dtype=np.float64)
        r = None if radius is None else np.ascontiguousarray(radius, dtype=np.float64)
        self._lib.nbs_data_set(self.handle, len(mass), _pd(pos), _pd(vel), _pd(mass), _pd(r))

    def get_count(self) -> int:
        return self._lib.nbs_data_count(self.handle)

    def bodies(self):
        n = self.get_count()
        pos = np.empty((n, 3)); vel = np.empty((n, 3)); m = np.empty(n); r = np.empty(n)
        self._lib.nbs_data_get(self.handle, _pd(pos), _pd(vel), _pd(m), _pd(r))
        return pos, vel, m, r

    def resize(self, n: int) -> None:
        """n bodies at the origin with zero mass (nbody_data::resize)."""
        self._lib.nbs_data_resize(self.handle, n)

    def copy(self) -> "nbody_data":
        """A new nbody_data with the same bodies, time and step."""
        d = nbody_data()
        pos, vel, m, r = self.bodies()
        d.set_bodies(pos, vel, m, r)
        d.set_time(self.get_time())
        d.set_step(self.get_step())
        return d

    def colors(self) -> np.ndarray:
        """RGBA float colours [n][4]."""
        c = np.empty((self.get_count(), 4), np.float32)
        self._lib.nbs_data_get_color(self.handle, c.ctypes.data_as(C.POINTER(C.c_float)))
        return c

    def get_time(self) -> float:
        return self._lib.nbs_data_time(self.handle)

    def get_step(self) -> int:
        return self._lib.nbs_data_step(self.handle)

    def set_time(self, t: float) -> None:
        self._lib.nbs_data_set_time(self.handle, t)

    def set_step(self, s: int) -> None:
        self._lib.nbs_data_set_step(self.handle, s)

    def advise_time(self, dt: float) -> None:
        self._lib.nbs_data_advise_time(self.handle, dt)

    def get_box_size(self) -> int:
        return self._lib.nbs_data_box_size(self.handle)

    def set_check_list(self, s: str) -> None:
        self._lib.nbs_data_set_check_list(self.handle, s.encode())

    def set_quiet(self, q: bool = True) -> None:
        self._lib.nbs_data_set_quiet(self.handle, 1 if q else 0)

    def print_statistics(self, engine: Optional["nbody_engine"]) -> None:
        self._lib.nbs_data_print_statistics(self.handle, engine.handle if engine else None)

    def statistics(self, engine: Optional["nbody_engine"] = None) -> np.ndarray:
        out = np.zeros(12)
        self._lib.nbs_data_statistics(self.handle, engine.handle if engine else None, _pd(out))
        return out

    def errors(self) -> np.ndarray:
        out = np.zeros(4)
        self._lib.nbs_data_errors(self.handle, _pd(out))
        return out

    def is_equal(self, other: "nbody_data", eps: float) -> bool:
        return bool(self._lib.nbs_data_is_equal(self.handle, other.handle, eps))


class memory:
    """An engine buffer (nbody_engine::memory); size in bytes."""

    def __init__(self, handle, owner=None):
        self.handle = handle
        self.owner = owner

    def size(self) -> int:
        return native.host_lib().nbs_memory_size(self.handle)

    def device_ptr(self):
        return native.host_lib().nbs_memory_device_ptr(self.handle)


class fake_memory(memory):
    """A buffer of a type no engine knows (test/engine/test_nbody_engine.cpp:748-759)."""

    def __init__(self, nbytes: int = 0):
        super().__init__(native.host_lib().nbs_memory_fake_create(nbytes))

    def __del__(self):
        if self.handle:
            native.host_lib().nbs_memory_fake_destroy(self.handle)
            self.handle = None


def _h(m: Optional[memory]):
    return None if m is None else m.handle


def _harr(ms: Sequence[Optional[memory]]):
    arr = (C.c_void_p * max(1, len(ms)))()
    for i, m in enumerate(ms):
        arr[i] = _h(m)
    return arr


def _carr(c: Optional[Sequence[float]]):
    if c is None:
        return None, None
    a = np.ascontiguousarray(c, dtype=np.float64)
    return a, _pd(a)


class nbody_engine:
    """nbody_engine (nbody/nbody_engine.h:16-96)."""

    def __init__(self, handle, lib=None):
        self._lib = native.host_lib()
        self.handle = handle
        self._data = None
        # strong references: the collector clears weak references before it
        # runs finalisers, and a solver must be finalised before its engine
        self._solvers = []

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            # a solver frees its buffers through its engine: destroy attached
            # solvers first (garbage collection may finalise them in any order)
            for s in list(getattr(self, "_solvers", ())):
                s.destroy()
            self._solvers = []
            self._lib.nbs_engine_destroy(h)
            self.handle = None

    def destroy(self) -> None:
        self.__del__()

    def type_name(self) -> str:
        return self._lib.nbs_engine_type_name(self.handle).decode()

    def init(self, data: nbody_data) -> bool:
        self._data = data
        return bool(self._lib.nbs_engine_init(self.handle, data.handle))

    def get_data(self, data: nbody_data) -> None:
        self._lib.nbs_engine_get_data(self.handle, data.handle)

    def problem_size(self) -> int:
        return self._lib.nbs_engine_problem_size(self.handle)

    def get_y(self) -> memory:
        return memory(self._lib.nbs_engine_get_y(self.handle), self)

    def advise_time(self, dt: float) -> None:
        self._lib.nbs_engine_advise_time(self.handle, dt)

    def get_time(self) -> float:
        return self._lib.nbs_engine_get_time(self.handle)

    def set_time(self, t: float) -> None:
        self._lib.nbs_engine_set_time(self.handle, t)

    def get_step(self) -> int:
        return self._lib.nbs_engine_get_step(self.handle)

    def set_step(self, s: int) -> None:
        self._lib.nbs_engine_set_step(self.handle, s)

    def fcompute(self, t: float, y: memory, f: memory) -> None:
        self._lib.nbs_engine_fcompute(self.handle, t, _h(y), _h(f))

    def clamp(self, y: memory, b: float) -> None:
        self._lib.nbs_engine_clamp(self.handle, _h(y), b)

    def create_buffer(self, nbytes: int) -> Optional[memory]:
        h = self._lib.nbs_engine_create_buffer(self.handle, nbytes)
        return memory(h, self) if h else None

    def create_buffers(self, nbytes: int, count: int) -> list:
        out = []
        for _ in range(count):
            m = self.create_buffer(nbytes)
            if m is None:
                self.free_buffers(out)
                return []
            out.append(m)
        return out

    def free_buffer(self, m: Optional[memory]) -> None:
        if m is not None and m.handle:
            self._lib.nbs_engine_free_buffer(self.handle, m.handle)
            m.handle = None

    def free_buffers(self, ms: list) -> None:
        for m in ms:
            self.free_buffer(m)
        ms.clear()

    def read_buffer(self, src: memory, dst: Optional[np.ndarray] = None) -> np.ndarray:
        if dst is None:
            dst = np.empty(src.size() // 8, dtype=np.float64)
        self._lib.nbs_engine_read_buffer(self.handle, dst.ctypes.data_as(C.c_void_p) if dst.size else None,
                                         _h(src))
        return dst

    def write_buffer(self, dst: memory, src: Optional[np.ndarray]) -> None:
        if src is not None:
            src = np.ascontiguousarray(src, dtype=np.float64)
            if src.nbytes < dst.size():
                raise ValueError("host array smaller than the buffer")
        self._lib.nbs_engine_write_buffer(self.handle, _h(dst),
                                          None if src is None else src.ctypes.data_as(C.c_void_p))

    def copy_buffer(self, a: memory, b: memory) -> None:
        self._lib.nbs_engine_copy_buffer(self.handle, _h(a), _h(b))

    def fill_buffer(self, a: memory, value: float) -> None:
        self._lib.nbs_engine_fill_buffer(self.handle, _h(a), value)

    def fmadd_inplace(self, a: memory, b: memory, c: float) -> None:
        self._lib.nbs_engine_fmadd_inplace(self.handle, _h(a), _h(b), c)

    def fmadd(self, a: memory, b: memory, c: memory, d: float) -> None:
        self._lib.nbs_engine_fmadd(self.handle, _h(a), _h(b), _h(c), d)

    def fmaddn_inplace(self, a: memory, b: Sequence[memory], c, csize: int) -> None:
        arr, p = _carr(c)
        self._lib.nbs_engine_fmaddn_inplace(self.handle, _h(a), _harr(b), len(b), p, csize)

    def fmaddn_corr(self, a: memory, corr: memory, b: Sequence[memory], c, csize: int) -> None:
        arr, p = _carr(c)
        self._lib.nbs_engine_fmaddn_corr(self.handle, _h(a), _h(corr), _harr(b), len(b), p, csize)

    def fmaddn(self, a: memory, b: Optional[memory], c: Sequence[memory], d, dsize: int) -> None:
        arr, p = _carr(d)
        self._lib.nbs_engine_fmaddn(self.handle, _h(a), _h(b), _harr(c), len(c), p, dsize)

    def fmaxabs(self, a: memory, initial: float = 0.0) -> float:
        r = C.c_double(initial)
        self._lib.nbs_engine_fmaxabs(self.handle, _h(a), C.byref(r))
        return r.value

    def get_compute_count(self) -> int:
        return self._lib.nbs_engine_compute_count(self.handle)

    def print_info(self) -> None:
        self._lib.nbs_engine_print_info(self.handle)

    def has_error(self) -> bool:
        return bool(self._lib.nbs_engine_has_error(self.handle))

    def hip_context(self):
        return self._lib.nbs_engine_hip_context(self.handle)

    def hip_tree(self):
        return self._lib.nbs_engine_hip_tree(self.handle)

    def hip_neighbours(self):
        """The nbh_ah state of a `hip_ah` engine (None for other engines)."""
        return self._lib.nbs_engine_hip_neighbours(self.handle)


def nbody_create_engine(params: dict) -> Optional[nbody_engine]:
    """Engine factory; None for an unknown engine or an invalid parameter."""
    h = native.host_lib().nbs_engine_create(_params(params))
    return nbody_engine(h) if h else None


class nbody_solver:
    """nbody_solver (nbody/nbody_solver.h:12-41)."""

    def __init__(self, handle):
        self._lib = native.host_lib()
        self.handle = handle
        self._engine = None

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self._lib.nbs_solver_destroy(h)
            self.handle = None
        e = getattr(self, "_engine", None)
        if e is not None:
            e._solvers = [s for s in e._solvers if s is not self]
            self._engine = None

    def destroy(self) -> None:
        self.__del__()

    def type_name(self) -> str:
        return self._lib.nbs_solver_type_name(self.handle).decode()

    def set_engine(self, e: nbody_engine) -> None:
        if self._engine is not None and self._engine is not e:
            self._engine._solvers = [s for s in self._engine._solvers if s is not self]
        self._engine = e
        if all(s is not self for s in e._solvers):
            e._solvers.append(self)
        self._lib.nbs_solver_set_engine(self.handle, e.handle)

    def set_time_step(self, min_step: float, max_step: float) -> None:
        self._lib.nbs_solver_set_time_step(self.handle, min_step, max_step)

    def run(self, data: nbody_data, max_time: float, dump_dt: float = 0.0, check_dt: float = 0.0,
            stream: Optional["nbody_data_stream"] = None) -> int:
        return self._lib.nbs_solver_run(self.handle, data.handle, stream.handle if stream is not None else None,
                                        max_time, dump_dt, check_dt)

    def advise(self, dt: float) -> None:
        self._lib.nbs_solver_advise(self.handle, dt)

    def print_info(self) -> None:
        self._lib.nbs_solver_print_info(self.handle)

    def rk_configure(self, max_recursion: int, subdivisions: int, refine_steps_count: int,
                     error_threshold: float) -> bool:
        return bool(self._lib.nbs_solver_rk_configure(self.handle, max_recursion, subdivisions,
                                                      refine_steps_count, error_threshold))

    def butcher_table(self):
        """(a[steps,steps], b1, b2, c, implicit, embedded) or None."""
        steps = C.c_size_t(0); imp = C.c_int(0); emb = C.c_int(0)
        if not self._lib.nbs_solver_rk_table(self.handle, C.byref(steps), C.byref(imp), C.byref(emb)):
            return None
        n = steps.value
        a = np.zeros((n, n)); b1 = np.zeros(n); b2 = np.zeros(n); c = np.zeros(n)
        self._lib.nbs_solver_rk_coeffs(self.handle, _pd(a), _pd(b1), _pd(b2), _pd(c))
        return a, b1, b2, c, bool(imp.value), bool(emb.value)

    def rk_substeps(self) -> int:
        return self._lib.nbs_solver_rk_substeps(self.handle)


def nbody_create_solver(params: dict) -> Optional[nbody_solver]:
    """Solver factory; None for an unknown solver or an invalid parameter."""
    h = native.host_lib().nbs_solver_create(_params(params))
    return nbody_solver(h) if h else None


class nbody_data_stream:
    """Frame-file writer (nbody/nbody_data_stream.h): <base>.idx, <base>K.dat,
    <base>.col, <base>.mass.  Methods return 0 on success, -1 on failure."""

    def __init__(self):
        self._lib = native.host_lib()
        self.handle = self._lib.nbs_stream_create()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self._lib.nbs_stream_destroy(h)
            self.handle = None

    def open(self, base: str, max_part_size: int = 1 << 30,
             append_to: Optional["nbody_data_stream_reader"] = None) -> int:
        return self._lib.nbs_stream_open(self.handle, base.encode(), max_part_size,
                                         append_to.handle if append_to is not None else None)

    def write(self, data: Optional[nbody_data]) -> int:
        return self._lib.nbs_stream_write(self.handle, data.handle if data is not None else None)

    def close(self) -> None:
        self._lib.nbs_stream_close(self.handle)


class nbody_data_stream_reader:
    """Frame-file reader (nbody/nbody_data_stream_reader.h)."""

    def __init__(self):
        self._lib = native.host_lib()
        self.handle = self._lib.nbs_stream_reader_create()

    def __del__(self):
        h = getattr(self, "handle", None)
        if h:
            self._lib.nbs_stream_reader_destroy(h)
            self.handle = None

    def load(self, base: str) -> int:
        return self._lib.nbs_stream_reader_load(self.handle, base.encode())

    def close(self) -> None:
        self._lib.nbs_stream_reader_close(self.handle)

    def get_frame_count(self) -> int:
        return self._lib.nbs_stream_reader_frame_count(self.handle)

    def get_steps_count(self) -> int:
        return self._lib.nbs_stream_reader_steps_count(self.handle)

    def get_max_time(self) -> float:
        return self._lib.nbs_stream_reader_max_time(self.handle)

    def seek(self, frame: int) -> int:
        return self._lib.nbs_stream_reader_seek(self.handle, frame)

    def get_current_frame(self) -> int:
        return self._lib.nbs_stream_reader_current_frame(self.handle)

    def get_body_count(self) -> int:
        return self._lib.nbs_stream_reader_body_count(self.handle)

    def get_coord_size(self) -> int:
        return self._lib.nbs_stream_reader_coord_size(self.handle)

    def get_box_size(self) -> float:
        return self._lib.nbs_stream_reader_box_size(self.handle)

    def get_last_file_n(self) -> int:
        return self._lib.nbs_stream_reader_last_file_n(self.handle)

    def read(self, data: Optional[nbody_data]) -> int:
        return self._lib.nbs_stream_reader_read(self.handle, data.handle if data is not None else None)
