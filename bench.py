#!/usr/bin/env python3
"""Direct-sum throughput of the `hip` engine on one MI355X.

Workload (the direct-sum configuration the metric is quoted on): N = 262144
bodies of a Plummer sphere (scale radius 1, total mass 1, seed 2), FP64 state,
FP32 force tiles with compensated accumulation, driven by the Dormand-Prince
5(4) solver (rkdp) with error_threshold 1e-9 and dt = 1e-3.  A step is one
rkdp advise(); the adaptive rule may split it into 8 substeps (one level, so a
step costs 7 or 63 force evaluations).  value = N^2 pair interactions per
force evaluation x evaluations / wall time of the timed steps, inputs resident
in HBM.

Also reported on the same line:
  roofline      the FP32 direct kernel's average duration (HIP events on the
                engine's stream) against the FP32 vector peak, 20 flop/pair;
  force_error   relative error of the FP32 forces vs the FP64 kernel at the
                final state (RMS and max over all bodies);
  cpu_baseline  the CPU restatement of the reference's `openmp` engine on the
                host cores, over a bounded sample of targets.
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time
from pathlib import Path

ROOT = Path(__file__).resolve().parent
sys.path.insert(0, str(ROOT))

METRIC = "pair-interactions/s (direct); BH steps/s at N=4M, force err ≤1e-3 vs FP64"
FLOP_PER_PAIR = 20.0
VALU_PER_PAIR = 13           # inner loop of k_direct_fp32 per pair (12 VALU + v_rsq_f32 with clamp)
PEAK_FP32_TFLOPS = 157.3


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(y, m, n, budget_s: float, threads: int) -> dict:
    """Pair rate of the CPU restatement (OpenMP over targets) on a sample of targets."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib
    ntgt = 256
    t = 0.0
    while True:
        idx = np.arange(ntgt, dtype=np.uint64)
        t0 = time.perf_counter()
        oracle_lib.accel_targets(y, m, idx, threads)
        t = time.perf_counter() - t0
        if t >= budget_s or ntgt >= n:
            break
        grow = max(2.0, min(16.0, budget_s / max(t, 1e-3)))
        ntgt = min(n, int(ntgt * grow))
    return {"value": ntgt * n / t, "unit": "pair-interactions/s", "cores": threads, "kind": "port",
            "sample": f"{ntgt} targets x {n} sources of the same Plummer state, FP64, "
                      f"{t:.1f} s on {threads} host threads"}


def _state_y(data):
    import numpy as np
    pos, vel, mass, _ = data.bodies()
    return np.ascontiguousarray(np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2]])), \
        mass


def bh_reference_accels(y, mass, idx) -> "np.ndarray":
    """FP64 direct-sum accelerations of the bodies idx (nbh_direct_sample: the
    oracle's operations and source order) of the state y."""
    import numpy as np

    from nbody_amd import native
    hip = native.hip_lib()
    n = len(mass)
    ctx = C.c_void_p()
    native.check(hip.nbh_ctx_create(0, C.byref(ctx)))
    ptrs = []

    def put(a):
        a = np.ascontiguousarray(a)
        ptr = C.c_void_p()
        native.check(hip.nbh_malloc(ctx, max(8, a.nbytes), C.byref(ptr)))
        native.check(hip.nbh_memcpy_h2d(ctx, ptr, a.ctypes.data_as(C.c_void_p), a.nbytes))
        ptrs.append(ptr)
        return ptr

    dy, dm, di = put(y), put(mass), put(np.asarray(idx, np.uint32))
    out = np.empty(3 * len(idx))
    da = put(out)
    native.check(hip.nbh_direct_sample(ctx, dy, dm, n, di, len(idx), da))
    native.check(hip.nbh_memcpy_d2h(ctx, out.ctypes.data_as(C.c_void_p), da, out.nbytes))
    for ptr in ptrs:
        native.check(hip.nbh_free(ctx, ptr))
    native.check(hip.nbh_ctx_destroy(ctx))
    return out.reshape(-1, 3)


def bh_cpu_baseline(y, mass, ratio: float, budget_s: float, threads: int) -> dict:
    """Euler steps/s of the CPU restatement of the reference's `simple_bh`
    engine (heap_stackless layout, nested_tree targets; nbody/nbody_engine_simple_bh.cpp:16-93)
    on the host cores: one full heap build of the state (OpenMP tasks, as
    nbody_space_heap::build_p) plus stackless walks of a growing sample of
    targets, extrapolated to all n targets."""
    import numpy as np
    sys.path.insert(0, str(ROOT / "tests"))
    import oracle_lib
    n = len(mass)
    h = oracle_lib.Heap(n)
    t0 = time.perf_counter()
    h.build_omp(y, mass, ratio, threads)
    t_build = time.perf_counter() - t0
    rng = np.random.default_rng(0)
    k = 1024
    while True:
        idx = rng.choice(n, k, replace=False)
        t0 = time.perf_counter()
        h.accel_targets(y, mass, idx, 2, threads)
        t_walk = time.perf_counter() - t0
        if t_build + t_walk >= budget_s or k >= n:
            break
        k = min(n, int(k * max(2.0, min(16.0, (budget_s - t_build) / max(t_walk, 1e-3)))))
    step_s = t_build + t_walk * n / k
    return {"value": 1.0 / step_s, "unit": "euler steps/s", "cores": threads, "kind": "port",
            "sample": f"theta={1.0 / ratio:g} heap_stackless walk, FP64: one full heap build of the N={n} state "
                      f"({t_build:.2f} s) + {k} sampled targets ({t_walk:.2f} s), extrapolated to N targets "
                      f"({step_s:.1f} s per force evaluation) on {threads} host threads"}


def bh_bench(n: int, steps: int, warmup: int, variants, samples: int = 8192, kind: str = "plummer",
             seed: int = 4, rk4_steps: int = 2, error_seeds=None) -> dict:
    """Barnes-Hut euler steps/s (one fcompute = tree build + group walk, plus one
    fmadd_inplace, as the reference bench's step) on a Plummer sphere of n bodies
    (seed `seed`), with the relative force error on `samples` random bodies
    against FP64 direct sums on the initial states of the seeds `error_seeds`
    (default: `seed` only).  variants: (label, ratio, kind[, engine keys]) --
    kind False: monopole walk; "quadrupole": monopole + quadrupole walk (keys
    bh_quad_near, bh_far_field); True: learned correction, trained first on
    another Plummer realisation of the same size (seed 7)."""
    import tempfile

    import numpy as np

    from nbody_amd import correction, native, nbody_create_engine, nbody_create_solver, nbody_data
    hip = native.hip_lib()
    error_seeds = list(error_seeds or [seed])
    if seed not in error_seeds:
        error_seeds.insert(0, seed)
    states = {}
    for sd in error_seeds:
        log(f"BH: generating {kind} N={n} seed {sd}")
        d = nbody_data()
        d.set_quiet(True)
        d.generate(kind, n, sd)
        states[sd] = (d,) + _state_y(d)
    data = states[seed][0]
    mass = states[seed][2]
    desc = {"plummer": "Plummer sphere (a=1, M=1)",
            "two_plummer": "two Plummer spheres of n/2 bodies (a=1, M=1 each), centres 10 apart, approaching at 0.5"}
    out = {"workload": f"Barnes-Hut (heap kd-tree, or Morton radix tree where a variant says so), FP32 forces / "
                       f"FP64 state, {desc.get(kind, kind)} N={n}, "
                       f"seed {seed}, euler dt=1e-4 (and rk4 steps/s), tree rebuilt every force evaluation; "
                       f"force error on {samples} sampled bodies of the initial states of seeds {error_seeds}",
           "n": n, "error_seeds": error_seeds, "runs": []}
    idx = np.sort(np.random.default_rng(0).choice(n, samples, replace=False)).astype(np.uint32)
    refs = {}
    for sd in error_seeds:
        refs[sd] = bh_reference_accels(states[sd][1], states[sd][2], idx)
    tmpdir = tempfile.mkdtemp(prefix="nbh_corr_")
    for label, ratio, learned, *extra in variants:
        params = {"engine": "hip_bh", "force_precision": "fp32", "bh_walk": "group",
                  "distance_to_node_radius_ratio": ratio, "tree_layout": "heap_stackless"}
        if extra:
            params.update(extra[0])
        if learned == "quadrupole":
            params["bh_multipole"] = "quadrupole"
            learned = False
        train_info = None
        if learned == "two_level":
            # the two-level local walk with the network's correction of the
            # super-groups' far fields, trained on another realisation (seed 7)
            params.update({"bh_multipole": "quadrupole", "bh_far_field": "local"})
            tdata = nbody_data()
            tdata.set_quiet(True)
            tdata.generate("plummer", n, 7)
            ty, tm = _state_y(tdata)
            near = float(params.get("bh_quad_near", 5.0))
            params["bh_quad_near"] = near
            log(f"BH: training the two-level correction at ratio {ratio}, near {near}")
            t0 = time.perf_counter()
            corr = correction.train_two_level(ty, tm, ratio, near_k=near, max_samples=2_000_000, epochs=40)
            train_info = dict(corr.stats, seconds=time.perf_counter() - t0, near_k=near,
                              train_state="Plummer seed 7")
            path = os.path.join(tmpdir, f"corr2_{ratio:g}.nbhmlp")
            correction.save(path, corr)
            params["learned_correction"] = path
        elif learned:
            tdata = nbody_data()
            tdata.set_quiet(True)
            tdata.generate("plummer", n, 7)
            ty, tm = _state_y(tdata)
            log(f"BH: training the correction at ratio {ratio}")
            t0 = time.perf_counter()
            # far cells beyond 8 group half-diagonals with radius / distance > 0.5 go
            # to the network, nearer ones get the per-body quadrupole (measured at 1M:
            # the most accurate and cheapest of the thresholds tried)
            corr = correction.train(ty, tm, ratio, rho_min=0.5, near_k=8.0, max_samples=2_000_000, epochs=20)
            train_info = dict(corr.stats, seconds=time.perf_counter() - t0, rho_min=corr.rho_min,
                              near_k=corr.near_k, train_state="Plummer seed 7")
            path = os.path.join(tmpdir, f"corr_{ratio:g}.nbhmlp")
            correction.save(path, corr)
            params["learned_correction"] = path
        e = nbody_create_engine(params)
        data.set_time(0.0)
        data.set_step(0)
        if e is None or not e.init(data):
            raise RuntimeError("hip_bh engine init failed")
        ctx = e.hip_context()
        s = nbody_create_solver({"solver": "euler", "max_step": 1e-4})
        s.set_engine(e)
        for _ in range(warmup):
            s.advise(1e-4)
        native.check(hip.nbh_timing_enable(ctx, 1))
        native.check(hip.nbh_timing_reset(ctx))
        native.check(hip.nbh_ctx_synchronize(ctx))
        t0 = time.perf_counter()
        for _ in range(steps):
            s.advise(1e-4)
        native.check(hip.nbh_ctx_synchronize(ctx))
        wall = time.perf_counter() - t0
        tb, cb, tw, cw = C.c_double(), C.c_uint64(), C.c_double(), C.c_uint64()
        native.check(hip.nbh_timing_read(ctx, b"tree_build", C.byref(tb), C.byref(cb)))
        native.check(hip.nbh_timing_read(ctx, b"bh_traverse", C.byref(tw), C.byref(cw)))
        tg, cg = C.c_double(), C.c_uint64()
        native.check(hip.nbh_timing_read(ctx, b"bh_group_walk", C.byref(tg), C.byref(cg)))
        native.check(hip.nbh_timing_enable(ctx, 0))
        cnt = (C.c_uint64 * 5)()
        native.check(hip.nbh_bh_walk_counters(ctx, e.hip_tree(), cnt))
        walk_counts = list(cnt)
        # force error on the initial state of every seed (same masses: m = 1/n)
        errs = {}
        f = e.create_buffer(48 * n)
        yb = e.create_buffer(48 * n)
        spills = walk_counts[2]
        for sd in error_seeds:
            e.write_buffer(yb, states[sd][1])
            e.fcompute(0.0, yb, f)
            native.check(hip.nbh_bh_walk_counters(ctx, e.hip_tree(), cnt))
            spills = max(spills, cnt[2])
            fa = e.read_buffer(f)[3 * n:].reshape(3, n).T[idx.astype(np.int64)]
            a0 = refs[sd]
            rel = np.linalg.norm(fa - a0, axis=1) / np.linalg.norm(a0, axis=1)
            errs[sd] = {"rms": float(np.sqrt(np.mean(rel ** 2))), "p99": float(np.percentile(rel, 99)),
                        "max": float(rel.max()), "samples": samples}
        e.free_buffers([f, yb])
        rk4_rate = None
        if rk4_steps > 0:
            s4 = nbody_create_solver({"solver": "rk4", "max_step": 1e-4})
            s4.set_engine(e)
            s4.advise(1e-4)
            native.check(hip.nbh_ctx_synchronize(ctx))
            t0 = time.perf_counter()
            for _ in range(rk4_steps):
                s4.advise(1e-4)
            native.check(hip.nbh_ctx_synchronize(ctx))
            rk4_rate = rk4_steps / (time.perf_counter() - t0)
            s4.destroy()
        walk_ms = tw.value / max(1, cw.value)
        run = {
            "variant": label, "ratio": ratio, "theta": 1.0 / ratio,
            "tree": params.get("bh_tree", "heap"),
            "multipole": params.get("bh_multipole", "monopole"),
            "quad_near": params.get("bh_quad_near", 3.0) if "bh_multipole" in params else None,
            "far_field": params.get("bh_far_field", "per_body") if "bh_multipole" in params else None,
            "steps_per_s": steps / wall,
            "ms_per_step": wall / steps * 1e3, "tree_build_ms": tb.value / max(1, cb.value),
            "walk_ms": walk_ms, "group_walk_ms": (tg.value / cg.value) if cg.value else None,
            "interactions_per_body": walk_counts[0] / n,
            "node_visits_per_body": walk_counts[1] / n,
            # node records read by the group walks: one 64-byte record per visit
            "node_GB_per_s": 64.0 * walk_counts[1] / (walk_ms * 1e-3) / 1e9,
            # nodes taken as monopoles because a walk stack was full, plus the
            # two-level walk's super-groups whose list overflowed (they walk in
            # one level); the largest count over the timed and the error runs
            "spills_or_fallbacks": int(spills),
            "rk4_steps_per_s": rk4_rate,
            "force_rel_err": errs[seed],
            "force_rel_err_by_seed": {str(k): v for k, v in errs.items()},
            "rms_max_over_seeds": max(v["rms"] for v in errs.values()),
            "p99_max_over_seeds": max(v["p99"] for v in errs.values()),
        }
        if train_info is not None or "bh_multipole" in params:
            # per 64 bodies (one heap group; Morton groups hold up to 64)
            run["far_cells_per_64_bodies"] = walk_counts[3] / (n / 64)
            run["near_cells_per_64_bodies"] = walk_counts[4] / (n / 64)
        if train_info is not None:
            run["training"] = train_info
        out["runs"].append(run)
        log(f"BH {label}: {steps / wall:.2f} steps/s, rms err {run['force_rel_err']['rms']:.2e} "
            f"(max over seeds {run['rms_max_over_seeds']:.2e}), spills/fallbacks {spills}")
        s.destroy()
        e.destroy()
    for name in os.listdir(tmpdir):
        os.remove(os.path.join(tmpdir, name))
    os.rmdir(tmpdir)
    ok = [r for r in out["runs"] if r["rms_max_over_seeds"] <= 1e-3]
    if ok:
        best = max(ok, key=lambda r: r["steps_per_s"])
        out["best_at_rms_1e-3"] = {"variant": best["variant"], "tree": best["tree"], "steps_per_s": best["steps_per_s"],
                                   "rms": best["force_rel_err"]["rms"],
                                   "rms_max_over_seeds": best["rms_max_over_seeds"],
                                   "walk_ms": best["walk_ms"], "tree_build_ms": best["tree_build_ms"]}
        base = [r for r in out["runs"] if r["variant"] == "plain theta=0.5"]
        if base:
            out["best_at_rms_1e-3"]["speedup_vs_plain_theta_0.5"] = best["steps_per_s"] / base[0]["steps_per_s"]
    # roofline of the dominant kernel of the fastest variant within RMS 1e-3: the group
    # walks of the two-level walk, their live HIP-event duration against the VALU
    # instructions and HBM bytes per launch that the counter runs in profiles/ measured
    # for this variant and state
    best_variant = out.get("best_at_rms_1e-3", {}).get("variant")
    rl = bh_roofline_from_profile(n, best_variant)
    if rl is not None:
        r = [x for x in out["runs"] if x["variant"] == rl["variant"] and x.get("group_walk_ms")]
        if r:
            ms = r[0]["group_walk_ms"]
            rate = rl["valu"] / 1024 / (ms * 1e-3) / (CLOCK_GHZ * 1e9)
            kern = "k_bh_group32<true, true, true, true>" if r[0]["tree"] == "morton" else \
                "k_bh_group32<true, true, true>"
            out["roofline"] = {
                "kernel": f"{kern} (group walks of the two-level walk)",
                "variant": rl["variant"], "bound": "valu",
                "achieved": rate, "peak": 0.5, "unit": "VALU wave instructions per SIMD per clock",
                "frac": rate / 0.5, "traffic": rl["traffic"],
                "hbm_GB_per_s": rl["traffic"] / (ms * 1e-3) / 1e9,
                "hbm_frac": rl["traffic"] / (ms * 1e-3) / 8e12, "avg_kernel_ms": ms,
                "source": rl["source"]}
    # equal-error comparison: for each plain monopole run, the fastest run of another
    # kind whose RMS and 99th-percentile errors are both no larger on every seed
    # (the RMS over the samples alone can swing with one body near a force
    # cancellation), and its speed relative to the plain run
    def plain(r):
        return r["multipole"] == "monopole" and r["tree"] == "heap" and not r["variant"].startswith("learned")

    learned_runs = [r for r in out["runs"] if r["variant"].startswith("learned")]
    base_runs = [r for r in out["runs"] if r["variant"] == "plain theta=0.5"]
    if learned_runs and base_runs:
        lr_, b_ = learned_runs[0], base_runs[0]
        same = [r for r in out["runs"] if not r["variant"].startswith("learned") and r["ratio"] == lr_["ratio"]
                and r["far_field"] == "local" and r["quad_near"] == lr_["quad_near"]]
        out["learned_gate"] = {
            "learned": lr_["variant"], "steps_per_s": lr_["steps_per_s"],
            "speedup_vs_plain_theta_0.5": lr_["steps_per_s"] / b_["steps_per_s"],
            "rms_max_over_seeds": lr_["rms_max_over_seeds"], "p99_max_over_seeds": lr_["p99_max_over_seeds"],
            "plain_theta_0.5_rms_max_over_seeds": b_["rms_max_over_seeds"],
            "plain_theta_0.5_p99_max_over_seeds": b_["p99_max_over_seeds"],
            "meets_gate": bool(lr_["steps_per_s"] >= 2.0 * b_["steps_per_s"]
                               and lr_["rms_max_over_seeds"] <= b_["rms_max_over_seeds"]
                               and lr_["p99_max_over_seeds"] <= b_["p99_max_over_seeds"]),
            "without_network": ({"variant": same[0]["variant"], "steps_per_s": same[0]["steps_per_s"],
                                 "rms_max_over_seeds": same[0]["rms_max_over_seeds"],
                                 "rms_by_seed": {k: v["rms"] for k, v in same[0]["force_rel_err_by_seed"].items()}}
                                if same else None),
            "rms_by_seed": {k: v["rms"] for k, v in lr_["force_rel_err_by_seed"].items()},
        }
        if same:
            # what the network itself does: the same walk with and without it
            out["learned_gate"]["network_lowers_error"] = bool(
                lr_["rms_max_over_seeds"] < same[0]["rms_max_over_seeds"])
    eq = []
    for p in out["runs"]:
        if not plain(p):
            continue
        cand = [r for r in out["runs"] if not plain(r)
                and all(r["force_rel_err_by_seed"][k]["rms"] <= p["force_rel_err_by_seed"][k]["rms"]
                        and r["force_rel_err_by_seed"][k]["p99"] <= p["force_rel_err_by_seed"][k]["p99"]
                        for k in p["force_rel_err_by_seed"])]
        if cand:
            b = max(cand, key=lambda r: r["steps_per_s"])
            eq.append({"plain": p["variant"], "plain_rms": p["force_rel_err"]["rms"],
                       "plain_p99": p["force_rel_err"]["p99"], "plain_rms_max_over_seeds": p["rms_max_over_seeds"],
                       "plain_steps_per_s": p["steps_per_s"], "fastest": b["variant"],
                       "rms": b["force_rel_err"]["rms"], "p99": b["force_rel_err"]["p99"],
                       "rms_max_over_seeds": b["rms_max_over_seeds"], "steps_per_s": b["steps_per_s"],
                       "speedup": b["steps_per_s"] / p["steps_per_s"]})
    if eq:
        out["at_equal_error"] = eq
    return out


def split_precision_run(n: int, steps: int, solvers=("rkfeagin10", "bs"), precisions=("fp64", "fp32", "fp32_corr"),
                        stream_base: str = "") -> dict:
    """Cold cube [-1,1]^3, v = 0, masses lognormal(0, 0.5) normalised to 1 (seed 5):
    `steps` steps of max_step 1e-3 (rkfeagin10 with one level of 8 adaptive
    substeps, bs with max_level 8) with FP32 force tiles, alone and with the FP64
    correction pass of the nearest sources (+ the compensated FP64 Butcher
    combination for rkfeagin10), against all-FP64; energy drift and the
    absolute total momentum / angular momentum of both runs (P0 = L0 = 0, so the
    reference's percentage errors are undefined for this state).

    stream_base: run in chunks through frame streams (nbody_data_stream):
    <stream_base>_<solver>_<precision> holds the initial state as frame 0 and the
    state at the end of every chunk; a call continues from the last frame and
    appends its own, and the drift is taken against frame 0."""
    import numpy as np

    from nbody_amd import nbody_create_engine, nbody_create_solver, nbody_data
    from nbody_amd import nbody_data_stream, nbody_data_stream_reader
    out = {"workload": f"cold cube N={n}, lognormal(0.5) masses, {steps} steps of 1e-3", "runs": []}

    def engine_for(d, prec):
        e = nbody_create_engine({"engine": "hip", "force_precision": prec})
        if e is None or not e.init(d):
            raise RuntimeError("hip engine init failed")
        return e

    def fresh():
        d = nbody_data()
        d.set_quiet(True)
        d.generate("cube", n, 5)
        d.set_check_list("PLEV")
        return d

    for solver in solvers:
        res = {"solver": solver}
        for prec in precisions:
            d = fresh()
            reader = None
            base = f"{stream_base}_{solver}_{prec}" if stream_base else ""
            if base and os.path.exists(base + ".idx"):
                reader = nbody_data_stream_reader()
                if reader.load(base) != 0 or reader.seek(0) != 0 or reader.read(d) != 0:
                    raise RuntimeError(f"cannot read frame 0 of {base}")
            e0_engine = engine_for(d, prec)
            st0 = d.statistics(e0_engine)
            e0_engine.destroy()
            stream = None
            if base:
                if reader is not None:
                    last = reader.get_frame_count() - 1
                    if reader.seek(last) != 0 or reader.read(d) != 0:
                        raise RuntimeError(f"cannot read frame {last} of {base}")
                stream = nbody_data_stream()
                if stream.open(base, 1 << 30, reader) != 0:
                    raise RuntimeError(f"cannot open stream {base}")
                if reader is None and stream.write(d) != 0:
                    raise RuntimeError(f"cannot write frame 0 of {base}")
            t_start, step_start = d.get_time(), d.get_step()
            e = engine_for(d, prec)
            params = {"solver": solver, "max_step": 1e-3}
            if solver == "bs":
                params["max_level"] = 8
            else:
                # one level of 8 substeps: with the default 8 levels the adaptive rule
                # can subdivide without bound on close pairs of this cold start
                params["max_recursion"] = 1
            if solver == "rkfeagin10" and prec != "fp64":
                params["correction"] = True
            s = nbody_create_solver(params)
            s.set_engine(e)
            t0 = time.perf_counter()
            for i in range(steps):
                s.advise(1e-3)
                if (i + 1) % max(1, steps // 20) == 0:
                    log(f"split precision {solver} {prec}: step {i + 1}/{steps}, "
                        f"{e.get_compute_count()} force evaluations, {time.perf_counter() - t0:.1f} s")
            e.get_data(d)
            st1 = d.statistics(e)
            if stream is not None:
                if stream.write(d) != 0:
                    raise RuntimeError(f"cannot append to {base}")
                stream.close()
            if reader is not None:
                reader.close()
            e0, e1 = st0[6] + st0[7], st1[6] + st1[7]
            res[prec] = {"energy_drift_rel": (e1 - e0) / abs(e0), "abs_P": float(np.linalg.norm(st1[:3])),
                         "abs_L": float(np.linalg.norm(st1[3:6])), "seconds": time.perf_counter() - t0,
                         "force_evaluations": e.get_compute_count(), "t_start": t_start, "t_end": d.get_time(),
                         "steps_start": step_start, "steps_end": d.get_step()}
            s.destroy()
            e.destroy()
        out["runs"].append(res)
    return out


def vector_ops_bench(n: int, reps: int = 20) -> dict:
    """HBM bandwidth of the engine's fused vector ops on a state of n bodies
    (6n doubles per buffer): nbh_fmaddn with k = 1 and k = 12 terms (the
    rkfeagin10 stage combinations), nbh_fmaddn_corr with k = 12, nbh_maxabs.
    Bytes are the algorithmic streams: fmaddn reads init + k sources and writes
    the result; fmaddn_corr also reads and writes the compensation; maxabs reads
    one buffer.  Device time from HIP events on the context's stream."""
    from nbody_amd import native
    hip = native.hip_lib()
    m = 6 * n
    ctx = C.c_void_p()
    native.check(hip.nbh_ctx_create(0, C.byref(ctx)))
    bufs = []

    def dev():
        ptr = C.c_void_p()
        native.check(hip.nbh_malloc(ctx, 8 * m, C.byref(ptr)))
        native.check(hip.nbh_fill(ctx, ptr, 1.0, m))
        bufs.append(ptr)
        return ptr

    kmax = 12
    a, init, corr = dev(), dev(), dev()
    srcs = [dev() for _ in range(kmax)]
    src_arr = (C.c_void_p * kmax)(*[p.value for p in srcs])
    coef = (C.c_double * kmax)(*([1e-3] * kmax))
    out = {"n": n, "doubles_per_buffer": m, "runs": []}
    cases = [("fmaddn k=1", lambda: hip.nbh_fmaddn(ctx, a, init, src_arr, coef, 1, m), 8.0 * m * 3, "fmaddn"),
             ("fmaddn k=12", lambda: hip.nbh_fmaddn(ctx, a, init, src_arr, coef, 12, m), 8.0 * m * 14, "fmaddn"),
             ("fmaddn_corr k=12", lambda: hip.nbh_fmaddn_corr(ctx, a, corr, src_arr, coef, 12, m), 8.0 * m * 16,
              "fmaddn")]
    for label, fn, nbytes, cls in cases:
        native.check(fn())
        native.check(hip.nbh_timing_enable(ctx, 1))
        native.check(hip.nbh_timing_reset(ctx))
        for _ in range(reps):
            native.check(fn())
        tot, cnt = C.c_double(), C.c_uint64()
        native.check(hip.nbh_timing_read(ctx, cls.encode(), C.byref(tot), C.byref(cnt)))
        native.check(hip.nbh_timing_enable(ctx, 0))
        ms = tot.value / max(1, cnt.value)
        out["runs"].append({"op": label, "bytes": nbytes, "ms": ms, "GB_per_s": nbytes / (ms * 1e-3) / 1e9,
                            "frac_of_8TBps": nbytes / (ms * 1e-3) / 8e12})
    val = C.c_double()
    t0 = time.perf_counter()
    for _ in range(reps):
        native.check(hip.nbh_maxabs(ctx, a, m, C.byref(val)))
    ms = (time.perf_counter() - t0) / reps * 1e3
    out["runs"].append({"op": "maxabs (incl. the 8-byte read-back)", "bytes": 8.0 * m, "ms": ms,
                        "GB_per_s": 8.0 * m / (ms * 1e-3) / 1e9, "frac_of_8TBps": 8.0 * m / (ms * 1e-3) / 8e12})
    for ptr in bufs:
        native.check(hip.nbh_free(ctx, ptr))
    native.check(hip.nbh_ctx_destroy(ctx))
    return out


def pmc_probe(n: int) -> int:
    """Launches for the hardware-counter runs: the FP32 direct kernel three times
    on the benchmark state, then a plain streaming kernel of known traffic
    (nbh_fmaddn a = b + c over 2^26 doubles: 1 GiB read, 512 MiB written) three
    times, so FETCH_SIZE / WRITE_SIZE of the first can be scaled by what the
    counters read for the second."""
    import numpy as np

    from nbody_amd import native, nbody_data
    hip = native.hip_lib()
    data = nbody_data()
    data.set_quiet(True)
    data.generate("plummer", n, 2)
    pos, vel, mass, _ = data.bodies()
    y = np.ascontiguousarray(np.concatenate([pos[:, 0], pos[:, 1], pos[:, 2], vel[:, 0], vel[:, 1], vel[:, 2]]))
    ctx = C.c_void_p()
    native.check(hip.nbh_ctx_create(0, C.byref(ctx)))
    bufs = []

    def dev(nbytes):
        ptr = C.c_void_p()
        native.check(hip.nbh_malloc(ctx, nbytes, C.byref(ptr)))
        bufs.append(ptr)
        return ptr

    dy, dm, df = dev(y.nbytes), dev(mass.nbytes), dev(y.nbytes)
    native.check(hip.nbh_memcpy_h2d(ctx, dy, y.ctypes.data_as(C.c_void_p), y.nbytes))
    native.check(hip.nbh_memcpy_h2d(ctx, dm, mass.ctypes.data_as(C.c_void_p), mass.nbytes))
    for _ in range(3):
        native.check(hip.nbh_fcompute_direct(ctx, dy, dm, df, n, 1, 0))
    native.check(hip.nbh_ctx_synchronize(ctx))
    log("direct launches done")
    for ptr in bufs:
        native.check(hip.nbh_free(ctx, ptr))
    native.check(hip.nbh_ctx_destroy(ctx))
    return copy_probe()


def traffic_from_profile(n: int):
    """HBM bytes per launch of the FP32 direct kernel at this N, from the counter
    summary in profiles/ (line 'traffic_bytes_per_launch N=<n>: <bytes>'), or None."""
    path = ROOT / "profiles" / f"direct_fp32_plummer_{n}_traffic.md"
    if not path.exists():
        return None
    key = f"traffic_bytes_per_launch N={n}:"
    for line in path.read_text().splitlines():
        if line.strip().startswith(key):
            return float(line.split(":", 1)[1].strip())
    return None


CLOCK_GHZ = 2.4
BH_COUNTERS = ("bh_two_level_walk_counters_plummer_{n}.md", "bh_morton_two_level_walk_plummer_{n}.md")
BH_COUNTER_VARIANT = {"theta=0.62 near=3": "quadrupole theta=0.62, local far field, near 3"}
for _th, _near in (("0.5", "3"), ("0.6", "3"), ("0.7", "3"), ("0.55", "2.5"), ("0.58", "2.5"), ("0.6", "2.5"),
                   ("0.5", "2"), ("0.55", "2"), ("0.58", "2"), ("0.6", "2")):
    BH_COUNTER_VARIANT[f"morton theta={_th} near={_near}"] = \
        f"morton tree: quadrupole theta={_th}, local far field, near {_near}"


def bh_roofline_from_profile(n: int, prefer=None):
    """VALU wave instructions and HBM bytes per launch of the two-level group walk
    at this N, from the counter summaries in profiles/ (lines
    'valu_wave_instr_per_launch <key>: <count>' and 'traffic_bytes_per_launch
    <key>: <bytes>'), with the bench variant they were measured on: the variant
    `prefer` where a summary holds it, else one on the same tree, else the first
    one; or None."""
    found = []
    for name in BH_COUNTERS:
        path = ROOT / "profiles" / name.format(n=n)
        if not path.exists():
            continue
        got = {}
        for line in path.read_text().splitlines():
            for tag, field in (("valu_wave_instr_per_launch ", "valu"), ("traffic_bytes_per_launch ", "traffic")):
                if line.startswith(tag) and ":" in line:
                    key, val = line[len(tag):].split(":", 1)
                    got.setdefault(key.strip(), {})[field] = float(val)
        found += [{"variant": BH_COUNTER_VARIANT[key], "valu": vals["valu"], "traffic": vals["traffic"],
                   "source": f"profiles/{path.name}"}
                  for key, vals in got.items() if key in BH_COUNTER_VARIANT and "valu" in vals and "traffic" in vals]
    for f in found:
        if f["variant"] == prefer:
            return f
    morton = prefer is not None and prefer.startswith("morton tree:")
    for f in found:
        if f["variant"].startswith("morton tree:") == morton:
            return f
    return found[0] if found else None


def bh_variants(learned: bool = True):
    """Barnes-Hut variants of the 4M Plummer block: (label, ratio, kind[, keys])."""
    local = {"bh_far_field": "local"}
    v = [("plain theta=0.5", 2.0, False), ("plain theta=0.8", 1.25, False),
         ("plain theta=0.25", 4.0, False),
         ("quadrupole theta=0.6, all cells per body", 1.0 / 0.6, "quadrupole", {"bh_quad_near": "inf"}),
         ("quadrupole theta=0.6", 1.0 / 0.6, "quadrupole"),
         ("quadrupole theta=0.8", 1.25, "quadrupole"),
         ("quadrupole theta=0.6, local far field", 1.0 / 0.6, "quadrupole", local),
         ("quadrupole theta=0.62, local far field, near 3", 1.0 / 0.62, "quadrupole",
          {**local, "bh_quad_near": 3.0}),
         ("quadrupole theta=0.65, local far field, near 4", 1.0 / 0.65, "quadrupole",
          {**local, "bh_quad_near": 4.0}),
         ("quadrupole theta=0.65, local far field, near 5", 1.0 / 0.65, "quadrupole",
          {**local, "bh_quad_near": 5.0}),
         ("quadrupole theta=0.7, local far field, near 4", 1.0 / 0.7, "quadrupole",
          {**local, "bh_quad_near": 4.0}),
         ("quadrupole theta=0.8, local far field, near 5", 1.25, "quadrupole",
          {**local, "bh_quad_near": 5.0}),
         ("quadrupole theta=0.9, local far field, near 4", 1.0 / 0.9, "quadrupole",
          {**local, "bh_quad_near": 4.0})]
    # the same two-level local walk on the Morton radix tree, whose cells are
    # compact where the heap's count-balanced cells span from the core outwards
    for th, near in ((0.5, 3), (0.6, 3), (0.7, 3), (0.55, 2.5), (0.58, 2.5), (0.6, 2.5), (0.5, 2), (0.55, 2),
                     (0.58, 2), (0.6, 2)):
        v.append((f"morton tree: quadrupole theta={th:g}, local far field, near {near:g}", 1.0 / th, "quadrupole",
                  {**local, "bh_quad_near": float(near), "bh_tree": "morton"}))
    if learned:
        v.append(("learned theta=0.8, two-level local walk, near 5", 1.25, "two_level", {"bh_quad_near": 5.0}))
    return v


def bh_probe(n: int, label: str) -> int:
    """Launches for the hardware-counter runs of one Barnes-Hut variant: three
    force evaluations (tree build + walk) of the 4M Plummer state (seed 4), then
    the calibration copy of pmc_probe."""
    from nbody_amd import native, nbody_create_engine, nbody_data
    hip = native.hip_lib()
    v = [x for x in bh_variants(False) if x[0] == label]
    if not v:
        log(f"unknown variant {label!r}; known: {[x[0] for x in bh_variants(False)]}")
        return 2
    lab, ratio, kind, *extra = v[0]
    params = {"engine": "hip_bh", "force_precision": "fp32", "bh_walk": "group",
              "distance_to_node_radius_ratio": ratio, "tree_layout": "heap_stackless"}
    if kind == "quadrupole":
        params["bh_multipole"] = "quadrupole"
    if extra:
        params.update(extra[0])
    data = nbody_data()
    data.set_quiet(True)
    data.generate("plummer", n, 4)
    e = nbody_create_engine(params)
    if e is None or not e.init(data):
        raise RuntimeError("hip_bh engine init failed")
    f = e.create_buffer(48 * n)
    for _ in range(3):
        e.fcompute(0.0, e.get_y(), f)
    native.check(hip.nbh_ctx_synchronize(e.hip_context()))
    e.free_buffer(f)
    e.destroy()
    log(f"BH launches done: {lab}")
    return copy_probe()


def copy_probe() -> int:
    """A plain streaming kernel of known traffic (nbh_fmaddn a = b + c over 2^26
    doubles: 1 GiB read, 512 MiB written), three times, to calibrate
    FETCH_SIZE / WRITE_SIZE."""
    from nbody_amd import native
    hip = native.hip_lib()
    ctx = C.c_void_p()
    native.check(hip.nbh_ctx_create(0, C.byref(ctx)))
    m = 1 << 26
    bufs = []
    for _ in range(3):
        ptr = C.c_void_p()
        native.check(hip.nbh_malloc(ctx, 8 * m, C.byref(ptr)))
        bufs.append(ptr)
    a, b, c = bufs
    native.check(hip.nbh_fill(ctx, b, 1.0, m))
    native.check(hip.nbh_fill(ctx, c, 2.0, m))
    srcs = (C.c_void_p * 1)(c.value)
    coef = (C.c_double * 1)(1.0)
    for _ in range(3):
        native.check(hip.nbh_fmaddn(ctx, a, b, srcs, coef, 1, m))
    native.check(hip.nbh_ctx_synchronize(ctx))
    log("copy launches done")
    for ptr in bufs:
        native.check(hip.nbh_free(ctx, ptr))
    native.check(hip.nbh_ctx_destroy(ctx))
    return 0


def main(argv=None) -> int:
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--n", type=int, default=1 << 18)
    ap.add_argument("--dt", type=float, default=1e-3)
    ap.add_argument("--precision", choices=["fp32", "fp64"], default="fp32")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-bh", action="store_true", help="skip the Barnes-Hut steps/s block")
    ap.add_argument("--bh-n", type=int, default=1 << 22)
    ap.add_argument("--bh-steps", type=int, default=5)
    ap.add_argument("--no-learned", action="store_true", help="skip the learned-correction Barnes-Hut run")
    ap.add_argument("--no-vector-ops", action="store_true", help="skip the vector-op bandwidth block")
    ap.add_argument("--no-two-plummer", action="store_true",
                    help="skip the colliding two-Plummer Barnes-Hut workload (N=2^20)")
    ap.add_argument("--split-precision", type=int, default=0, metavar="STEPS",
                    help="also run the cold-cube split-precision workload (N=65536) for STEPS steps")
    ap.add_argument("--split-solver", choices=["rkfeagin10", "bs"], action="append",
                    help="split-precision solvers (default: both)")
    ap.add_argument("--split-prec", choices=["fp64", "fp32", "fp32_corr"], action="append",
                    help="split-precision force precisions (default: all three)")
    ap.add_argument("--split-stream", default="", metavar="BASE",
                    help="run the split-precision workload in chunks through frame streams BASE_<solver>_<prec>")
    ap.add_argument("--pmc-probe", action="store_true",
                    help="only launch the kernels measured by the hardware-counter runs, then exit")
    ap.add_argument("--bh-probe", metavar="VARIANT",
                    help="only launch three force evaluations of this Barnes-Hut variant at --bh-n (for the "
                         "hardware-counter runs) and the calibration copy, then exit")
    a = ap.parse_args(argv)

    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print("bench.py: this build runs on one GPU; multi-GPU sharding is not part of it "
              "(use --gpus 1)", file=sys.stderr)
        return 2

    import numpy as np

    import nbody_amd
    from nbody_amd import native, nbody_create_engine, nbody_create_solver, nbody_data

    if nbody_amd.device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        return 2
    if a.pmc_probe:
        return pmc_probe(a.n)
    if a.bh_probe:
        return bh_probe(a.bh_n, a.bh_probe)
    hip = native.hip_lib()
    n = a.n
    log(f"generating Plummer N={n}")
    data = nbody_data()
    data.set_quiet(True)
    data.generate("plummer", n, 2)
    e = nbody_create_engine({"engine": "hip", "force_precision": a.precision})
    if e is None or not e.init(data):
        print(f"bench.py: engine init failed: {nbody_amd.log_last()}", file=sys.stderr)
        return 1
    s = nbody_create_solver({"solver": "rkdp", "error_threshold": 1e-9, "max_step": a.dt,
                             "max_recursion": 1, "substep_subdivisions": 8})
    s.set_engine(e)
    ctx = e.hip_context()
    native.check(hip.nbh_timing_enable(ctx, 1))

    for i in range(a.warmup):
        s.advise(a.dt)
        native.check(hip.nbh_ctx_synchronize(ctx))
        log(f"warmup step {i + 1}/{a.warmup} done")
    native.check(hip.nbh_timing_reset(ctx))
    cc0 = e.get_compute_count()
    sub0 = s.rk_substeps()
    native.check(hip.nbh_ctx_synchronize(ctx))
    t0 = time.perf_counter()
    for i in range(a.steps):
        s.advise(a.dt)
        if (i + 1) % max(1, a.steps // 5) == 0:
            log(f"step {i + 1}/{a.steps}")
    native.check(hip.nbh_ctx_synchronize(ctx))
    wall = time.perf_counter() - t0
    if e.has_error():
        print(f"bench.py: HIP error during the run: {nbody_amd.log_last()}", file=sys.stderr)
        return 1
    fcomputes = e.get_compute_count() - cc0
    substeps = s.rk_substeps() - sub0
    pairs = float(fcomputes) * n * n
    value = pairs / wall

    kname = b"direct_fp32" if a.precision == "fp32" else b"direct_fp64"
    tot = C.c_double(0)
    cnt = C.c_uint64(0)
    native.check(hip.nbh_timing_read(ctx, kname, C.byref(tot), C.byref(cnt)))
    avg_ms = tot.value / max(1, cnt.value)
    achieved = FLOP_PER_PAIR * n * n / (avg_ms * 1e-3) / 1e12
    peak = PEAK_FP32_TFLOPS if a.precision == "fp32" else 78.6
    native.check(hip.nbh_timing_enable(ctx, 0))

    # force error of the final state: FP32 tiles vs the FP64 kernel
    log("force error check")
    y = e.get_y()
    f32m = e.create_buffer(48 * n)
    f64m = e.create_buffer(48 * n)
    yp = y.device_ptr()
    yh = e.read_buffer(y)
    _, _, mass, _ = data.bodies()
    dm = e.create_buffer(8 * n)
    e.write_buffer(dm, mass)
    massp = dm.device_ptr()
    native.check(hip.nbh_fcompute_direct(ctx, yp, massp, f32m.device_ptr(), n, 1, 0))
    native.check(hip.nbh_fcompute_direct(ctx, yp, massp, f64m.device_ptr(), n, 0, 256))
    f32 = e.read_buffer(f32m)[3 * n:].reshape(3, n)
    f64 = e.read_buffer(f64m)[3 * n:].reshape(3, n)
    rel = np.linalg.norm(f32 - f64, axis=0) / np.linalg.norm(f64, axis=0)
    e.free_buffers([f32m, f64m, dm])

    out = {
        "metric": METRIC,
        "value": value,
        "unit": "pair-interactions/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": wall / a.steps * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": a.precision,
        "data": "synthetic",
        "config": {
            "workload": f"direct all-pairs, Plummer N={n} (a=1, M=1, seed 2), rkdp error_threshold 1e-9, "
                        f"dt={a.dt:g}, max_recursion 1 x 8 substeps; {a.precision.upper()} force tiles, FP64 state",
            "n": n,
            "fcompute_per_step": fcomputes / a.steps,
            "accepted_substeps": substeps,
            "parallelism": "single-gpu",
        },
        "roofline": {
            "bound": "valu",
            "kernel": kname.decode(),
            "achieved": achieved,
            "peak": peak,
            "unit": "TFLOP/s",
            "frac": achieved / peak,
            "traffic": traffic_from_profile(n) if a.precision == "fp32" else None,
            "avg_kernel_ms": avg_ms,
            "launches": cnt.value,
            "pairs_per_s_kernel": n * n / (avg_ms * 1e-3),
            # issued instructions: 12 VALU + 1 v_rsq_f32 per pair against the
            # 157.3e12 / 2 lane-instructions/s behind the FP32 peak, and as wave
            # instructions per SIMD per clock (256 CUs x 4 SIMDs, 2.4 GHz)
            "valu_instr_per_pair": VALU_PER_PAIR,
            "valu_issue_frac": VALU_PER_PAIR * n * n / (avg_ms * 1e-3) / (peak * 1e12 / 2),
            "valu_wave_instr_per_simd_clk": VALU_PER_PAIR * n * n / 64 / (avg_ms * 1e-3) / (1024 * 2.4e9),
        },
        "force_error": {"rms_rel": float(np.sqrt(np.mean(rel ** 2))), "max_rel": float(rel.max()),
                        "vs": "FP64 direct kernel (bit-exact with the CPU oracle)"},
    }
    s.destroy()
    e.destroy()
    if not a.no_cpu_baseline:
        threads = min(16, os.cpu_count() or 1)
        log(f"cpu baseline ({threads} threads, ~{a.cpu_seconds:g} s)")
        out["cpu_baseline"] = cpu_baseline(yh, mass, n, a.cpu_seconds, threads)
    if not a.no_bh:
        variants = bh_variants(not a.no_learned)
        bh = bh_bench(a.bh_n, a.bh_steps, 1, variants, error_seeds=[4, 5, 6])
        if not a.no_cpu_baseline:
            threads = min(16, os.cpu_count() or 1)
            log(f"BH cpu baseline ({threads} threads, ~{a.cpu_seconds:g} s)")
            from nbody_amd import nbody_data as _nd
            d4 = _nd()
            d4.set_quiet(True)
            d4.generate("plummer", a.bh_n, 4)
            y4, m4 = _state_y(d4)
            bh["cpu_baseline"] = bh_cpu_baseline(y4, m4, 2.0, a.cpu_seconds, threads)
        out["barnes_hut"] = bh
    if not a.no_bh and not a.no_two_plummer:
        out["two_plummer"] = bh_bench(1 << 20, 3, 1, [("monopole theta=0.5, group walk", 2.0, False),
                                                      ("quadrupole theta=0.6", 1.0 / 0.6, "quadrupole"),
                                                      ("quadrupole theta=0.6, local far field", 1.0 / 0.6,
                                                       "quadrupole", {"bh_far_field": "local"}),
                                                      ("morton tree: quadrupole theta=0.58, local far field, near 3",
                                                       1.0 / 0.58, "quadrupole",
                                                       {"bh_far_field": "local", "bh_quad_near": 3.0,
                                                        "bh_tree": "morton"})],
                                      kind="two_plummer", seed=3, rk4_steps=3)
    if not a.no_vector_ops:
        log("vector ops")
        out["vector_ops"] = vector_ops_bench(1 << 22)
    if a.split_precision > 0:
        out["split_precision"] = split_precision_run(
            1 << 16, a.split_precision, tuple(a.split_solver or ("rkfeagin10", "bs")),
            tuple(a.split_prec or ("fp64", "fp32", "fp32_corr")), a.split_stream)
    print(json.dumps(out), flush=True)
    return 0


if __name__ == "__main__":
    sys.exit(main())
