"""Object lifetime of the Python wrappers: a solver frees its buffers through
its engine, so garbage collection must finalise it before the engine, whatever
order the collector picks for a cycle that holds both."""
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]

SCRIPT = r"""
import gc, sys
sys.path.insert(0, {tests!r}); sys.path.insert(0, {root!r})
import oracle_lib
from nbody_amd import nbody_create_solver, nbody_data
for order in range(4):
    d = nbody_data(); d.set_quiet(True); d.generate("ball", 64, 1)
    e = oracle_lib.simple_engine(1)
    assert e.init(d)
    s = nbody_create_solver({{"solver": "rk4", "max_step": 1e-3}})
    s.set_engine(e)
    s.advise(1e-3)
    cycle = [e, s, d] if order % 2 else [s, e, d]
    cycle.append(cycle)
    del d, e, s, cycle
    gc.collect()
print("ok")
"""


def test_cycle_with_engine_and_solver_is_collected_safely():
    code = SCRIPT.format(tests=str(ROOT / "tests"), root=str(ROOT))
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr[-2000:]
    assert r.stdout.strip().endswith("ok")
