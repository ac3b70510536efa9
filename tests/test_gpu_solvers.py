"""The reference's solver known-answer tests (test/solver/test_nbody_solver.cpp)
with the `hip` engine, and whole trajectories of the `hip` engine against the
oracle `simple` engine under the same solvers (bit-exact in FP64).
"""
import numpy as np
import pytest

import oracle_lib
from nbody_amd import nbody_create_engine
from solver_cases import CASES, golden, max_abs_diff, run_case

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("name,params", CASES, ids=[c[0] for c in CASES])
def test_solver_golden_on_hip(name, params):
    e = nbody_create_engine({"engine": "hip"})
    data, s = run_case(e, params)
    exp = golden(name)
    assert exp.is_equal(data, 1e-12), f"{name}: max |diff| = {max_abs_diff(exp, data):.3e}"
    assert not e.has_error()
    s.destroy()
    e.destroy()


@pytest.mark.parametrize("name,params", [c for c in CASES if c[0] in
                                         ("rk4", "rkdp-corr", "rkfeagin10", "bulirsch-stoer", "adams5-corr")],
                         ids=lambda c: c if isinstance(c, str) else "")
def test_solver_trajectory_bit_exact(name, params):
    """Same solver, same inputs: the hip engine's FP64 state equals the oracle's bit for bit."""
    eh = nbody_create_engine({"engine": "hip"})
    eo = oracle_lib.simple_engine(1)
    dh, sh = run_case(eh, params)
    do, so = run_case(eo, params)
    ph, vh, _, _ = dh.bodies()
    po, vo, _, _ = do.bodies()
    assert np.array_equal(ph, po) and np.array_equal(vh, vo), \
        f"max |diff| {max(np.max(np.abs(ph - po)), np.max(np.abs(vh - vo))):.3e}"
    assert eh.get_compute_count() == eo.get_compute_count()
    for x in (sh, so, eh, eo):
        x.destroy()
