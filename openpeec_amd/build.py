"""Build recipe for libpeec_mi355x.so (C host code + gfx950 HIP kernels), in-tree.

    python -m openpeec_amd.build          # compile everything (always rebuilds)

C host sources are compiled with gcc (-ffp-contract=off: the reference's x86-64 build never
fuses multiply-adds, and the host-side geometry/packing must match it bit for bit); HIP sources
with hipcc for gfx950 only.  Objects go to openpeec_amd/_build/, the shared library next to
this file.  The checker builds under oracle/ are driven from __graft_entry__.build().
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libpeec_mi355x.so"

ARCH = "gfx950"
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
INC = [f"-I{ROOT / 'include'}", f"-I{CSRC}"]

C_FLAGS = ["-std=gnu11", "-O2", "-fPIC", "-Wall", "-Wno-unused-function", "-ffp-contract=off", "-fopenmp"]
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd: list[str]) -> None:
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {' '.join(cmd)}")
    noise = "unknown target feature"  # hipcc's host pass on the packed-fp32 flag (harmless)
    err = "\n".join(l for l in res.stderr.splitlines() if noise not in l and "packed-fp32" not in l)
    if err.strip():
        sys.stderr.write(err + "\n")


def _obj(src: Path) -> Path:
    return BUILD / (src.parent.name + "_" + src.stem + ".o")


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    c_srcs = sorted((CSRC / "host").glob("*.c"))
    hip_srcs = sorted((CSRC / "hip").glob("*.hip"))
    jobs = []
    for s in c_srcs:
        jobs.append(["gcc", *C_FLAGS, *INC, "-c", str(s), "-o", str(_obj(s))])
    for s in hip_srcs:
        jobs.append([HIPCC, *HIP_FLAGS, *INC, "-c", str(s), "-o", str(_obj(s))])
    nproc = min(16, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=nproc) as ex:
        futs = [ex.submit(_run, j) for j in jobs]
        for f in futs:
            f.result()
    objs = [str(_obj(s)) for s in c_srcs + hip_srcs]
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-Wl,-Bsymbolic", "-o", str(tmp), *objs, "-lrocblas", "-lgomp", "-lm"])
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB}")
    return LIB


if __name__ == "__main__":
    build(verbose=True)
