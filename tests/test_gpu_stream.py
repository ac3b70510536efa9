"""Frame dumps of a run on the `hip` engine (asynchronous snapshots: a staging
copy on the engine's stream, then a pinned copy on a second stream while the
next steps run) against the same run on the oracle's `simple` engine, and a
run resumed from its stream against the uninterrupted run."""
import numpy as np
import pytest

import oracle_lib
from nbody_amd import (nbody_create_engine, nbody_create_solver, nbody_data, nbody_data_stream,
                       nbody_data_stream_reader)

pytestmark = pytest.mark.gpu


def _data(n=1024, seed=1):
    d = nbody_data()
    d.set_quiet(True)
    d.generate("ball", n, seed)
    return d


def _run(engine, base, t1, data=None, append_to=None, max_part=1 << 30):
    data = data if data is not None else _data()
    assert engine.init(data)
    s = nbody_create_solver({"solver": "rk4", "max_step": 1e-3})
    s.set_engine(engine)
    st = nbody_data_stream()
    assert st.open(base, max_part, append_to=append_to) == 0
    assert s.run(data, t1, 5e-3, 0.0, stream=st) == 0
    st.close()
    assert not engine.has_error()
    s.destroy()
    return data


def _frames(base):
    r = nbody_data_stream_reader()
    assert r.load(base) == 0
    d = nbody_data()
    d.resize(r.get_body_count())
    out = []
    for _ in range(r.get_frame_count()):
        assert r.read(d) == 0
        pos, vel, _, _ = d.bodies()
        out.append((d.get_step(), d.get_time(), pos.copy(), vel.copy()))
    return out


def test_hip_dumps_equal_oracle_dumps(tmp_path):
    """rk4, dt 1e-3, a frame every 5e-3 to t = 0.03 (7 frames over several
    data files): the `hip` FP64 frames equal the oracle's bit for bit."""
    eh = nbody_create_engine({"engine": "hip"})
    _run(eh, str(tmp_path / "hip" / "s"), 0.03, max_part=100_000)
    eo = oracle_lib.simple_engine(1)
    _run(eo, str(tmp_path / "cpu" / "s"), 0.03, max_part=100_000)
    fh, fo = _frames(str(tmp_path / "hip" / "s")), _frames(str(tmp_path / "cpu" / "s"))
    assert len(fh) == len(fo) == 7
    for a, b in zip(fh, fo):
        assert a[0] == b[0] and a[1] == b[1]
        assert np.array_equal(a[2], b[2]) and np.array_equal(a[3], b[3])
    assert not np.array_equal(fh[0][2], fh[-1][2])
    eh.destroy()
    eo.destroy()


def test_hip_resume_matches_uninterrupted_run(tmp_path):
    """Dump to t = 0.0155 (last frame at step 15), resume from that frame to
    t = 0.0305: the same last frame, index and frame files as one run to
    t = 0.0305 (both take 31 steps; the index keeps 6 significant digits of
    each time, so the bounds sit between steps)."""
    full, part = str(tmp_path / "full" / "s"), str(tmp_path / "part" / "s")
    e = nbody_create_engine({"engine": "hip"})
    final = _run(e, full, 0.0305)
    e.destroy()
    e = nbody_create_engine({"engine": "hip"})
    _run(e, part, 0.0155)
    e.destroy()
    r = nbody_data_stream_reader()
    assert r.load(part) == 0
    assert r.seek(r.get_frame_count() - 1) == 0
    d = nbody_data()
    d.set_quiet(True)
    d.resize(r.get_body_count())
    assert r.read(d) == 0
    e = nbody_create_engine({"engine": "hip"})
    assert d.get_step() == 15
    resumed = _run(e, part, 0.0305, data=d, append_to=r)
    e.destroy()
    assert resumed.is_equal(final, 0.0)
    with open(full + ".idx") as a, open(part + ".idx") as b:
        assert a.read() == b.read()
    with open(full + "0.dat", "rb") as a, open(part + "0.dat", "rb") as b:
        assert a.read() == b.read()


def test_cli_run_and_resume(tmp_path):
    """nbody-simulation with --engine=hip: a run with dumps, then --resume of
    its stream to a later time; the resumed stream equals one run to that time."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent.parent / "nbody_amd" / "nbody-simulation"
    common = ["--engine=hip", "--solver=rk4", "--max_step=0.001", "--stars_count=64", "--box_size=100",
              "--dump_step=0.005", "--check_step=0.01", "--verbose=1"]
    full, part = tmp_path / "full" / "s", tmp_path / "part" / "s"
    r = subprocess.run([str(exe), *common, "--max_time=0.0305", f"--output={full}"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "Engine: nbody_engine_hip" in r.stdout
    r = subprocess.run([str(exe), *common, "--max_time=0.0155", f"--output={part}"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    r = subprocess.run([str(exe), *common, "--max_time=0.0305", f"--resume={part}"], capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert (Path(str(full) + ".idx")).read_text() == (Path(str(part) + ".idx")).read_text()
    assert (Path(str(full) + "0.dat")).read_bytes() == (Path(str(part) + "0.dat")).read_bytes()
    assert len(_frames(str(full))) == 7


def test_bench_front_end_tree_ratio():
    """nbody-bench --bench=gpu_tree_ratio (the reference's gpu_tree_ratio sweep,
    test/bench/main.cpp:148-216) on a small state: one row per hip_bh variant,
    one positive time per step per ratio."""
    import subprocess
    from pathlib import Path
    exe = Path(__file__).resolve().parent.parent / "nbody_amd" / "nbody-bench"
    r = subprocess.run([str(exe), "--bench=gpu_tree_ratio", "--stars_count=16384", "--values=2,4",
                        "--max_time=0.03"], capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stderr[-2000:]
    lines = [line for line in r.stdout.splitlines() if not line.startswith("%")]
    assert lines[0] == "time" and lines[1].split() == ["name", "2", "4"]
    rows = {line.split()[0]: [float(v) for v in line.split()[1:]] for line in lines[2:]}
    assert set(rows) == {"hip_bh+heap+nested", "hip_bh+heap_stackless", "hip_bh+group", "hip_bh+two_level"}
    assert all(len(v) == 2 and all(t > 0 for t in v) for v in rows.values())
