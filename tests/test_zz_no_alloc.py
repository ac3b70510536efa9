"""nbh_fcompute_direct_noalloc through the C ABI: the same bits as
nbh_fcompute_direct on two calls in a row (the second with fewer bodies, so
stale scratch from the first would show), for every precision; sizes beyond
its fixed scratch are refused before any launch."""
import numpy as np
import pytest

import oracle_lib
from hip_capi import Ctx
from nbody_amd import native, nbody_data

pytestmark = pytest.mark.gpu

NBH_FP64, NBH_FP32, NBH_FP32_CORR, NBH_FP64_BLOCK = 0, 1, 2, 3
NBH_ERR_INVALID = 1
MAX_BODIES = 131072


@pytest.fixture(scope="module")
def ctx():
    c = Ctx(0)
    yield c
    c.close()


def _state(kind: str, n: int, seed: int):
    d = nbody_data()
    d.set_quiet(True)
    d.generate(kind, n, seed)
    return oracle_lib.state_from_data(d)


def _both(ctx, y, m, prec):
    n = len(m)
    dy, dm = ctx.put(y), ctx.put(m)
    fa, fb = ctx.alloc(6 * n), ctx.alloc(6 * n)
    native.check(ctx.lib.nbh_fcompute_direct(ctx.h, dy.ptr, dm.ptr, fa.ptr, n, prec, 0), "fcompute_direct")
    native.check(ctx.lib.nbh_fcompute_direct_noalloc(ctx.h, dy.ptr, dm.ptr, fb.ptr, n, prec, 0),
                 "fcompute_direct_noalloc")
    return fa.read(), fb.read()


@pytest.mark.parametrize("prec", [NBH_FP32, NBH_FP32_CORR, NBH_FP64, NBH_FP64_BLOCK])
def test_noalloc_matches_direct_twice(ctx, prec):
    # the largest size first, then a smaller one
    cases = [("plummer", MAX_BODIES, 3), ("ball", 5001, 1)]
    if prec in (NBH_FP64, NBH_FP64_BLOCK):
        cases = [("plummer", 20000, 3), ("ball", 5001, 1)]    # FP64 costs ~30x more per pair
    for kind, n, seed in cases:
        print(f"[noalloc] precision {prec} {kind} n={n}", flush=True)
        y, m = _state(kind, n, seed)
        fa, fb = _both(ctx, y, m, prec)
        assert np.array_equal(fa, fb), f"{kind} n={n}: noalloc result differs"
        assert np.all(np.isfinite(fb))


def test_noalloc_fp32_near_oracle(ctx):
    # the FP32 result is a force field, not a stale buffer: sampled targets vs the FP64 oracle
    n = 40000
    y, m = _state("plummer", n, 5)
    _, fb = _both(ctx, y, m, NBH_FP32)
    idx = np.arange(0, n, 97, dtype=np.uint64)
    a0 = oracle_lib.accel_targets(y, m, idx, 8)
    a = fb[3 * n:].reshape(3, n).T[idx.astype(np.int64)]
    rel = np.linalg.norm(a - a0, axis=1) / np.linalg.norm(a0, axis=1)
    assert np.sqrt(np.mean(rel ** 2)) < 1e-5 and rel.max() < 1e-3


def test_noalloc_refuses_too_many_bodies(ctx):
    n = MAX_BODIES + 1
    dy, dm, df = ctx.alloc(6 * n), ctx.alloc(n), ctx.alloc(6 * n)
    st = ctx.lib.nbh_fcompute_direct_noalloc(ctx.h, dy.ptr, dm.ptr, df.ptr, n, NBH_FP32, 0)
    assert st == NBH_ERR_INVALID
    assert "exceeds" in ctx.lib.nbh_last_error().decode()
