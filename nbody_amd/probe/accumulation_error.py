"""FP32 error of two ways to sum the direct-sum tile terms, emulated on the CPU.

For the targets of one 1024-body work-group of the FP32 direct kernel (Morton
order, Plummer sphere, N = 2^18, seed 2) against all sources, in tiles of 256:

  dx form    the kernel's arithmetic: per pair dx = x_s' - x_t' (scaled FP32
             offsets from the work-group origin), q = min(rsq(r'^2), 1),
             p += dx * (m' q) (q q), tile sums folded into a Kahan pair;
  mfma form  the contraction a matrix-core kernel would do: S += q^3 (m' x_s'),
             Sm += q^3 m' per tile (sources in Morton order, the self pair masked
             in the work-group's own tiles), Kahan across tiles, then
             a = S - x_t' Sm in FP64;

and prints the relative error of each against the FP64 sum (RMS and max over the
1024 targets).  FP32 rounding is emulated with numpy float32 (fma as one FP64
product and sum rounded once; rsq as a correctly rounded 1/sqrt).

Usage: python -m nbody_amd.probe.accumulation_error [work-group index, default 128]
(about a minute per work-group).
"""
from __future__ import annotations

import sys

import numpy as np

from nbody_amd import nbody_data

F32 = np.float32
LEN, MASS = 1e4, 1e8          # the kernel's scaled units: x' = 1e4 x, m' = 1e8 m
TILE, GROUP = 256, 1024


def morton_order(pos: np.ndarray) -> np.ndarray:
    lo = pos.min(0)
    ext = pos.max(0) - lo
    q = np.clip(((pos - lo) * (1024.0 / ext)).astype(np.int64), 0, 1023).astype(np.uint64)

    def spread(v):
        out = np.zeros_like(v)
        for b in range(10):
            out |= ((v >> np.uint64(b)) & np.uint64(1)) << np.uint64(3 * b)
        return out

    key = spread(q[:, 0]) | (spread(q[:, 1]) << np.uint64(1)) | (spread(q[:, 2]) << np.uint64(2))
    return np.argsort(key, kind="stable")


def fma(a, b, c):
    return (a.astype(np.float64) * b + c).astype(F32)


def rsq_clamped(r2):
    with np.errstate(divide="ignore"):
        return np.minimum((1.0 / np.sqrt(r2.astype(np.float64))).astype(F32), F32(1))


def kahan(s, c, v):
    y = (v - c).astype(F32)
    t = (s + y).astype(F32)
    return t, ((t - s).astype(F32) - y).astype(F32)


def exact(pos, m, tg):
    x = pos[tg]
    a = np.zeros((len(tg), 3))
    for s in range(0, len(m), 4096):
        d = pos[None, s:s + 4096, :] - x[:, None, :]
        r2 = np.maximum((d * d).sum(-1), 1e-8)
        a += (d * (m[None, s:s + 4096] / (r2 * np.sqrt(r2)))[..., None]).sum(1)
    return a


def dx_form(pos, m, tg, origin):
    xt = ((pos[tg] - pos[origin]) * LEN).astype(F32)
    xs = ((pos - pos[origin]) * LEN).astype(F32)
    ms = (m * MASS).astype(F32)
    S = np.zeros((len(tg), 3), F32)
    C = np.zeros_like(S)
    for base in range(0, len(m), TILE):
        p = np.zeros_like(S)
        for j in range(base, base + TILE):
            d = (xs[j] - xt).astype(F32)
            r2 = fma(d[:, 2], d[:, 2], fma(d[:, 1], d[:, 1], (d[:, 0] * d[:, 0]).astype(F32)))
            q = rsq_clamped(r2)
            w = ((ms[j] * q).astype(F32) * (q * q).astype(F32)).astype(F32)
            p = fma(d, w[:, None], p)
        S, C = kahan(S, C, p)
    return S.astype(np.float64) - C


def mfma_form(pos, m, perm, tg, origin, tbase):
    xt = ((pos[tg] - pos[origin]) * LEN).astype(F32)
    xs = ((pos[perm] - pos[origin]) * LEN).astype(F32)
    ms = (m[perm] * MASS).astype(F32)
    B = np.concatenate([(ms[:, None] * xs).astype(F32), ms[:, None]], 1)
    S = np.zeros((len(tg), 4), F32)
    C = np.zeros_like(S)
    for base in range(0, len(m), TILE):
        own = tbase <= base < tbase + GROUP
        p = np.zeros_like(S)
        for j in range(base, base + TILE):
            d = (xs[j] - xt).astype(F32)
            r2 = fma(d[:, 2], d[:, 2], fma(d[:, 1], d[:, 1], (d[:, 0] * d[:, 0]).astype(F32)))
            q = rsq_clamped(r2)
            q3 = ((q * q).astype(F32) * q).astype(F32)
            if own:
                q3 = np.where(r2 == 0, F32(0), q3)
            p = fma(q3[:, None], B[j][None, :], p)
        S, C = kahan(S, C, p)
    S64 = S.astype(np.float64) - C
    return S64[:, :3] - xt.astype(np.float64) * S64[:, 3:4]


def main() -> None:
    g = int(sys.argv[1]) if len(sys.argv) > 1 else 128
    d = nbody_data()
    d.set_quiet(True)
    d.generate("plummer", 1 << 18, 2)
    pos, _, m, _ = d.bodies()
    perm = morton_order(pos)
    tbase = g * GROUP
    tg = perm[tbase:tbase + GROUP]
    ref = exact(pos, m, tg)

    def err(a):
        e = np.linalg.norm(a - ref, axis=1) / np.linalg.norm(ref, axis=1)
        return f"RMS {np.sqrt((e ** 2).mean()):.3e}  max {e.max():.3e}"

    print(f"work-group {g} (Morton positions {tbase} .. {tbase + GROUP - 1}), N = {len(m)}")
    print("dx form,   origin at the first target :", err(dx_form(pos, m, tg, perm[tbase])), flush=True)
    print("mfma form, origin at the first target :", err(mfma_form(pos, m, perm, tg, perm[tbase], tbase)), flush=True)
    print("mfma form, origin at the middle target:",
          err(mfma_form(pos, m, perm, tg, perm[tbase + GROUP // 2], tbase)), flush=True)


if __name__ == "__main__":
    main()
