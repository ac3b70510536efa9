"""End-to-end workloads of the force path on the GPU against the oracle.

* Unit ball, N = 1024 (m = 1/N, velocities N(0, 0.3^2)), rk4 dt = 1e-3 for 100
  steps: `hip` (FP64) against the oracle's `simple` engine, state bit-exact and
  P/L/E/V every 10 steps (nbody_data::print_statistics, nbody/nbody_data.cpp:57-102).
* Cold cube with lognormal masses (the split-precision workload, scaled down):
  rkfeagin10 and Bulirsch-Stoer (max_level 8) with FP32 force tiles and the
  solvers' compensated FP64 combination against all-FP64; the FP64 run is
  checked bit-exact against the oracle, the FP32 run by its energy drift.
* Plummer sphere, N = 2^18 (seed 2), the direct-sum throughput workload at its
  full size: FP32 tiles against the oracle's FP64 direct sum on 8192 sampled
  targets, and two rkdp steps (error_threshold 1e-9, one recursion level) with
  FP32 tiles against the same run in FP64.
* Two Plummer spheres approaching (the Barnes-Hut workload, scaled down),
  theta = 0.5, rk4: the tree is rebuilt each step; node order and moments of
  the final state bit-exact with the oracle's heap, forces near FP64 direct.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib
from nbody_amd import native, nbody_create_engine, nbody_create_solver, nbody_data

pytestmark = pytest.mark.gpu


def _data(kind, n, seed):
    d = nbody_data()
    d.set_quiet(True)
    d.generate(kind, n, seed)
    return d


def _stats_rel(a, b):
    """Largest difference of two P/L/E/V vectors relative to the largest entry."""
    return np.max(np.abs(a - b)) / np.max(np.abs(b))


def test_unit_ball_rk4_plev_every_10_steps():
    n, dt = 1024, 1e-3
    dh, do = _data("ball", n, 1), _data("ball", n, 1)
    dh.set_check_list("PLEV")
    do.set_check_list("PLEV")
    _, _, m, _ = dh.bodies()
    assert np.allclose(m, 1.0 / n)
    eh = nbody_create_engine({"engine": "hip"})
    eo = oracle_lib.simple_engine(1)
    assert eh.init(dh) and eo.init(do)
    sh = nbody_create_solver({"solver": "rk4", "max_step": dt})
    so = nbody_create_solver({"solver": "rk4", "max_step": dt})
    sh.set_engine(eh)
    so.set_engine(eo)
    s0 = None
    for step in range(1, 101):
        sh.advise(dt)
        so.advise(dt)
        if step % 10 == 0:
            eh.get_data(dh)
            eo.get_data(do)
            ph, vh, _, _ = dh.bodies()
            po, vo, _, _ = do.bodies()
            assert np.array_equal(ph, po) and np.array_equal(vh, vo), step
            st_gpu = dh.statistics(eh)          # device compensated sums
            st_cpu = do.statistics(None)        # host sums of the oracle state
            assert _stats_rel(st_gpu, st_cpu) < 1e-13, (step, st_gpu, st_cpu)
            if s0 is None:
                s0 = st_cpu
    # conservation over the run (rk4, 100 steps of 1e-3): momentum to rounding,
    # energy to the integrator's error on close pairs (measured 1.9e-5 relative)
    assert np.max(np.abs(st_cpu[:3] - s0[:3])) < 1e-14
    e0, e1 = s0[6] + s0[7], st_cpu[6] + st_cpu[7]
    assert abs(e1 - e0) < 1e-4 * abs(e0)
    assert not eh.has_error()
    for x in (sh, so, eh, eo):
        x.destroy()


def _energy(d, e):
    d.set_check_list("PLEV")
    st = d.statistics(e)
    return st[6] + st[7]


@pytest.mark.parametrize("solver", [
    {"solver": "rkfeagin10", "max_step": 1e-3},
    {"solver": "bs", "max_step": 1e-3, "max_level": 8},
], ids=["rkfeagin10", "bs"])
def test_cold_cube_split_precision(solver):
    n, dt, steps = 4096, 1e-3, 20 if solver["solver"] == "rkfeagin10" else 8
    runs = {}
    counts = {}
    for label, prec, corr in (("fp64", "fp64", False), ("fp32", "fp32", True), ("fp32_corr", "fp32_corr", True),
                              ("oracle", None, False)):
        d = _data("cube", n, 5)
        _, v, m, _ = d.bodies()
        assert np.all(v == 0) and abs(m.sum() - 1.0) < 1e-12
        e = oracle_lib.simple_engine(0) if prec is None else nbody_create_engine({"engine": "hip",
                                                                                  "force_precision": prec})
        assert e.init(d)
        params = dict(solver)
        if corr and solver["solver"] == "rkfeagin10":
            params["correction"] = True
        s = nbody_create_solver(params)
        s.set_engine(e)
        e_start = _energy(d, e if prec is not None else None)
        for _ in range(steps):
            s.advise(dt)
        e.get_data(d)
        counts[label] = e.get_compute_count()
        runs[label] = (d, e_start, _energy(d, e if prec is not None else None))
        s.destroy()
        e.destroy()
    p64, v64, _, _ = runs["fp64"][0].bodies()
    po, vo, _, _ = runs["oracle"][0].bodies()
    assert np.array_equal(p64, po) and np.array_equal(v64, vo)
    assert counts["fp64"] == counts["oracle"]
    drift64 = runs["fp64"][2] - runs["fp64"][1]
    drift32 = runs["fp32"][2] - runs["fp32"][1]
    driftc = runs["fp32_corr"][2] - runs["fp32_corr"][1]
    e0 = abs(runs["fp64"][1])
    p32, _, _, _ = runs["fp32"][0].bodies()
    pc, _, _, _ = runs["fp32_corr"][0].bodies()
    dev = np.linalg.norm(p32 - p64, axis=1)
    devc = np.linalg.norm(pc - p64, axis=1)
    print(f"\n[split precision] {solver['solver']}: energy drift / |E0|: FP64 {drift64 / e0:.3e}, "
          f"FP32 tiles {drift32 / e0:.3e}, FP32 + FP64 correction pass {driftc / e0:.3e}; "
          f"position deviation from FP64 max {dev.max():.2e} / {devc.max():.2e}, p99 "
          f"{np.percentile(dev, 99):.2e} / {np.percentile(devc, 99):.2e}; force evaluations {counts}")
    # the split-precision runs follow the FP64 energy history to the FP32
    # rounding of close pairs (measured for rkfeagin10, 20 steps: 6.9e-6 of |E0|)
    for drift in (drift32, driftc):
        assert abs(drift - drift64) < 3e-5 * e0, (drift, drift64, e0)
    # most bodies agree to ~1e-12; close pairs carry the FP32 offset rounding
    # (measured max 3.5e-6 for rkfeagin10, 20 steps)
    for d in (dev, devc):
        assert np.percentile(d, 99) < 1e-9 and d.max() < 1e-4, (np.percentile(d, 99), d.max())
    if solver["solver"] == "rkfeagin10":
        # rkfeagin10 adapts its steps to its embedded error estimate.  The
        # correction pass takes the close pairs' FP32 rounding out, so the run
        # keeps the FP64 run's step decisions (measured: 1428 force evaluations
        # each over 20 steps; the FP32 tiles alone took 1020) and stays at least
        # as close to FP64 as the tiles do
        assert counts["fp32_corr"] == counts["fp64"]
        assert devc.max() <= dev.max() and abs(driftc - drift64) <= abs(drift32 - drift64)


def test_plummer_2e18_fp32_tiles():
    """The direct-sum workload at full size: N = 2^18 Plummer bodies (seed 2).

    (a) Force evaluation as the reference's engine harness does it
    (test/engine/test_nbody_engine.cpp:472-550: the derivative buffer is
    filled with a sentinel, fcompute, read back): the `hip` engine's FP32 tiles
    against the oracle's FP64 direct sum on 8192 sampled targets, RMS and p99
    relative error below 1e-5, maximum below 1e-3; the velocity half is an exact copy.
    (b) Two rkdp steps of dt = 1e-3 (error_threshold 1e-9, max_recursion 1)
    with FP32 tiles against the same run in FP64: the FP32 noise floor of the
    error estimate subdivides every step into 8 substeps (7 + 8 * 7 = 63 force
    evaluations); the states stay within the force error times the elapsed time."""
    n, seed, dt = 1 << 18, 2, 1e-3
    d = _data("plummer", n, seed)
    pos, vel, mass, _ = d.bodies()
    y = np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2]])
    e = nbody_create_engine({"engine": "hip", "force_precision": "fp32"})
    assert e.init(d)
    f = e.create_buffer(48 * n)
    e.fill_buffer(f, -1e10)
    e.fcompute(0.0, e.get_y(), f)
    out = e.read_buffer(f)
    e.free_buffer(f)
    assert np.array_equal(out[:3 * n], y[3 * n:])
    idx = np.random.default_rng(2).choice(n, 8192, replace=False)
    a0 = oracle_lib.accel_targets(y, mass, idx.astype(np.uint64), 0)
    a = out[3 * n:].reshape(3, n).T[idx]
    rel = np.linalg.norm(a - a0, axis=1) / np.linalg.norm(a0, axis=1)
    rms, p99 = float(np.sqrt(np.mean(rel ** 2))), float(np.percentile(rel, 99))
    print(f"\n[plummer 2^18] FP32 tiles vs oracle on 8192 targets: rms {rms:.2e} p99 {p99:.2e} max {rel.max():.2e}")
    assert rms < 1e-5 and p99 < 1e-5 and rel.max() < 1e-3
    e.destroy()

    finals, counts = {}, {}
    for prec in ("fp32", "fp64"):
        d = _data("plummer", n, seed)
        e = nbody_create_engine({"engine": "hip", "force_precision": prec})
        assert e.init(d)
        s = nbody_create_solver({"solver": "rkdp", "max_step": dt, "error_threshold": 1e-9, "max_recursion": 1})
        s.set_engine(e)
        per_step = []
        for _ in range(2):
            c0 = e.get_compute_count()
            s.advise(dt)
            per_step.append(e.get_compute_count() - c0)
        assert not e.has_error()
        e.get_data(d)
        finals[prec] = d.bodies()
        counts[prec] = per_step
        s.destroy()
        e.destroy()
    (p32, v32, _, _), (p64, v64, _, _) = finals["fp32"], finals["fp64"]
    dv = np.linalg.norm(v32 - v64, axis=1)
    dx = np.linalg.norm(p32 - p64, axis=1)
    print(f"[plummer 2^18] rkdp 2 steps, force evaluations per step FP32 {counts['fp32']} FP64 {counts['fp64']}; "
          f"|dv| max {dv.max():.2e} p99 {np.percentile(dv, 99):.2e}, |dx| max {dx.max():.2e}")
    assert counts["fp32"] == [63, 63]
    assert all(c in (7, 63) for c in counts["fp64"])
    # |a| < 1 for these bodies; the FP32 force error (max 1e-3, RMS 1e-5 relative)
    # over t = 2e-3 bounds the velocity deviation, t times that the positions
    assert dv.max() < 2e-6 and np.percentile(dv, 99) < 1e-7 and dx.max() < 4e-9


def test_two_plummer_bh_rk4():
    n, dt = 1 << 16, 1e-3
    d = _data("two_plummer", n, 3)
    e = nbody_create_engine({"engine": "hip_bh", "force_precision": "fp32", "tree_layout": "heap_stackless",
                             "distance_to_node_radius_ratio": 2.0})
    assert e.init(d)
    s = nbody_create_solver({"solver": "rk4", "max_step": dt})
    s.set_engine(e)
    for _ in range(5):
        s.advise(dt)
    e.get_data(d)
    assert not e.has_error()
    pos, vel, mass, _ = d.bodies()
    y = np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2]])
    # the tree the engine built for the last force evaluation's positions is
    # rebuilt here from the final state by both sides: same layout, bit for bit
    h = oracle_lib.Heap(n)
    h.build(y, mass, 2.0)
    o = h.arrays()
    f = e.create_buffer(48 * n)
    e.fcompute(0.0, e.get_y(), f)                # rebuilds the tree at the final state
    lib = native.hip_lib()
    s2 = 2 * n
    body = np.empty(s2, np.int32)
    cm = np.empty(3 * s2)
    native.check(lib.nbh_tree_get(e.hip_context(), e.hip_tree(), body.ctypes.data_as(C.c_void_p),
                                  cm.ctypes.data_as(C.c_void_p), None, None, None, None))
    assert np.array_equal(body[n:].astype(np.int64), o["body_n"][n:])
    assert np.array_equal(cm.reshape(-1, 3)[1:], o["cm"][1:])
    # forces near the FP64 direct sum
    a = e.read_buffer(f)[3 * n:].reshape(3, n).T
    idx = np.random.default_rng(0).choice(n, 512, replace=False)
    a0 = oracle_lib.accel_targets(y, mass, idx.astype(np.uint64), 0)
    rel = np.linalg.norm(a[idx] - a0, axis=1) / np.linalg.norm(a0, axis=1)
    assert np.sqrt(np.mean(rel ** 2)) < 2e-2
    e.free_buffer(f)
    s.destroy()
    e.destroy()
