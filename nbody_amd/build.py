"""Build the native libraries in-tree.

    nbody_amd/libnbody_hip.so   gfx950 kernels + C ABI (include/nbody_hip.h), hipcc
    nbody_amd/libnbody.so       host layer: engines, solvers, data (include/nbody_host.h), g++
    nbody_amd/nbody-simulation  command-line front end over libnbody.so, g++
    oracle/liboracle.so         CPU restatement used by the tests as the checker, gcc
    oracle/liboracle_engine.so  the restatement behind the nbody_engine interface (tests only)

Usage: python -m nbody_amd.build [--jobs N] [--no-oracle]
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
PKG = ROOT / "nbody_amd"
CSRC = PKG / "csrc"
HOST = PKG / "host"
INC = ROOT / "include"
ORACLE = ROOT / "oracle"
BUILD = PKG / "_build"

ARCH = os.environ.get("NBODY_OFFLOAD_ARCH", "gfx950")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")

# -packed-fp32-ops: no v_pk_*_f32 instructions are generated.
# -ffp-contract=off: every FP64 operation rounds where the source writes it.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", f"-I{INC}", f"-I{CSRC}",
]
HOST_FLAGS = ["-O2", "-fPIC", "-std=c++17", "-ffp-contract=off", "-Wall", "-Wextra",
              "-Wno-unused-parameter", f"-I{INC}", f"-I{HOST}", "-I/opt/rocm/include",
              "-D__HIP_PLATFORM_AMD__"]
ORACLE_FLAGS = ["-O3", "-fPIC", "-ffp-contract=off", "-fopenmp", "-Wall", "-Wextra", "-std=c11"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")
    out = "\n".join(l for l in r.stdout.splitlines()
                    if "packed-fp32-ops" not in l and "not a recognized feature" not in l
                    and l.strip() and not l.startswith("1 warning"))
    if out.strip():
        sys.stderr.write(out + "\n")


def _compile(job):
    cmd, src = job
    _run(cmd)
    return src


def build(jobs: int = 8, oracle: bool = True, verbose: bool = False) -> dict:
    BUILD.mkdir(parents=True, exist_ok=True)
    hip_srcs = sorted(CSRC.glob("*.hip"))
    host_srcs = sorted(HOST.glob("*.cpp"))
    compile_jobs = []
    hip_objs, host_objs = [], []
    for s in hip_srcs:
        o = BUILD / (s.stem + ".hip.o")
        hip_objs.append(o)
        compile_jobs.append(([HIPCC, *HIP_FLAGS, "-c", str(s), "-o", str(o)], s))
    for s in host_srcs:
        o = BUILD / (s.stem + ".o")
        host_objs.append(o)
        compile_jobs.append((["g++", *HOST_FLAGS, "-c", str(s), "-o", str(o)], s))
    if oracle:
        o = BUILD / "nbody_oracle.o"
        compile_jobs.append((["gcc", *ORACLE_FLAGS, f"-I{ORACLE}", "-c", str(ORACLE / "nbody_oracle.c"),
                              "-o", str(o)], ORACLE / "nbody_oracle.c"))
        o2 = BUILD / "oracle_engine.o"
        compile_jobs.append((["g++", *HOST_FLAGS, f"-I{ORACLE}", "-c", str(ORACLE / "oracle_engine.cpp"),
                              "-o", str(o2)], ORACLE / "oracle_engine.cpp"))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for src in ex.map(_compile, compile_jobs):
            if verbose:
                print("compiled", src.relative_to(ROOT))
    lib_hip = PKG / "libnbody_hip.so"
    lib_host = PKG / "libnbody.so"
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(lib_hip) + ".tmp",
          *map(str, hip_objs)])
    os.replace(str(lib_hip) + ".tmp", lib_hip)
    _run(["g++", "-shared", "-fPIC", "-o", str(lib_host) + ".tmp", *map(str, host_objs),
          f"-L{PKG}", "-lnbody_hip", "-Wl,-rpath,$ORIGIN"])
    os.replace(str(lib_host) + ".tmp", lib_host)
    out = {"libnbody_hip": lib_hip, "libnbody": lib_host}
    # command-line front ends, one executable per source
    for src in sorted((PKG / "cli").glob("*.cpp")):
        exe = PKG / src.stem.replace("_", "-")
        _run(["g++", *[f for f in HOST_FLAGS if f != "-fPIC"], "-o", str(exe) + ".tmp", str(src), f"-L{PKG}",
              "-lnbody", "-lnbody_hip", "-lpthread", "-Wl,-rpath,$ORIGIN"])
        os.replace(str(exe) + ".tmp", exe)
        out[src.stem] = exe
    if oracle:
        lib_or = ORACLE / "liboracle.so"
        _run(["gcc", "-shared", "-fopenmp", "-o", str(lib_or) + ".tmp", str(BUILD / "nbody_oracle.o"), "-lm"])
        os.replace(str(lib_or) + ".tmp", lib_or)
        lib_oe = ORACLE / "liboracle_engine.so"
        _run(["g++", "-shared", "-fopenmp", "-o", str(lib_oe) + ".tmp", str(BUILD / "oracle_engine.o"),
              str(BUILD / "nbody_oracle.o"), f"-L{PKG}", "-lnbody", "-lnbody_hip",
              "-Wl,-rpath,$ORIGIN/../nbody_amd"])
        os.replace(str(lib_oe) + ".tmp", lib_oe)
        out["liboracle"] = lib_or
        out["liboracle_engine"] = lib_oe
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=min(8, os.cpu_count() or 8))
    ap.add_argument("--no-oracle", action="store_true")
    ap.add_argument("-v", "--verbose", action="store_true")
    a = ap.parse_args(argv)
    libs = build(a.jobs, oracle=not a.no_oracle, verbose=a.verbose)
    for k, v in libs.items():
        print(f"{k}: {v.relative_to(ROOT)}")
    return 0


if __name__ == "__main__":
    sys.exit(main())
