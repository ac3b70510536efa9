"""The command-line front end (nbody_amd/nbody-simulation, the reference's
nbody-simulation, test/main/main.cpp:12-123) and its argument parser
(nbody/nbody_arg_parser.cpp:5-28), on the CPU: parsing, and the setup
failures that end the program before any GPU work."""
import ctypes as C
import subprocess
from pathlib import Path

from nbody_amd import native

EXE = Path(__file__).resolve().parent.parent / "nbody_amd" / "nbody-simulation"


def _parse(*args):
    argv = (C.c_char_p * (len(args) + 1))(b"prog", *[a.encode() for a in args])
    lib = native.host_lib()
    buf = C.create_string_buffer(4096)
    n = lib.nbs_parse_arguments(len(args) + 1, argv, buf, len(buf))
    text = buf.value.decode()
    assert len(text) == n
    return dict(kv.split("=", 1) for kv in text.split(";")) if text else {}


def test_parse_arguments():
    lib = native.host_lib()
    n0 = lib.nbs_log_count()
    got = _parse("--engine=hip_bh", "positional", "-x=1", "--verbose", "--a=b=c", "--max_time=0.5",
                 "--engine=hip", "--output=/tmp/x y")
    # the last value of a repeated key wins; arguments without "--" are skipped
    # silently, those without exactly one '=' with a log line
    assert got == {"engine": "hip", "max_time": "0.5", "output": "/tmp/x y"}
    assert lib.nbs_log_count() - n0 == 2
    assert _parse() == {}
    assert _parse("--solver=rk4", "--max_step=1e-3") == {"solver": "rk4", "max_step": "1e-3"}


def _run(*args):
    return subprocess.run([str(EXE), *args], capture_output=True, text=True, timeout=120)


def test_setup_failures(tmp_path):
    assert EXE.exists(), "front end not built: run python -m nbody_amd.build"
    r = _run("--engine=nonexistent", "--solver=rk4", f"--output={tmp_path}/s")
    assert r.returncode == 255 and "can't create engine" in r.stderr
    r = _run(f"--resume={tmp_path}/missing")
    assert r.returncode == 255 and "can't open stream to resume" in r.stderr
    r = _run(f"--initial_state={tmp_path}/missing.txt", "--initial_type=G1")
    assert r.returncode == 255 and "can't load initial state" in r.stderr


def test_bench_front_end_rejects_unknown_bench():
    exe = EXE.parent / "nbody-bench"
    assert exe.exists(), "front end not built: run python -m nbody_amd.build"
    r = subprocess.run([str(exe), "--bench=nonexistent"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 255 and "unknown bench" in r.stderr
