"""Build the in-tree native libraries of kmertreecounter_amd.

libkmc.so      hipcc --offload-arch=gfx950: the HIP engine (kernels + C ABI of
               include/kmc.h), the host FASTQ reader / output writers
               (include/kmc_host.h) and the synthetic data generator
               (include/kmc_synth.h)
kmc_tree       drop-in for the reference `Tree` command line, linked to libkmc.so

Everything is compiled every time build() runs (it is the "does it build"
check); nothing is installed outside the repository.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
LIB = PKG / "libkmc.so"
TREE = PKG / "kmc_tree"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = "gfx950"

HIP_SOURCES = ["kmc_kernels.hip", "kmc_fastq.hip", "kmc_engine.hip", "kmc_plan.hip", "kmc_learned.hip",
               "kmc_noalloc.hip"]
CXX_SOURCES = ["kmc_host.cpp", "synth.cpp"]

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++20", "-fPIC", "-fvisibility=hidden",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-result", f"-I{ROOT / 'include'}",
]
CXX_FLAGS = ["-O3", "-std=c++20", "-fPIC", "-fvisibility=hidden", "-Wall", "-march=x86-64-v2",
             f"-I{ROOT / 'include'}"]


def _run(cmd: list[str]) -> None:
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + r.stdout + r.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} ... {cmd[-1]}")


def build(verbose: bool = False) -> Path:
    BUILD.mkdir(exist_ok=True)
    objs = []
    jobs = []
    for src in HIP_SOURCES:
        p = CSRC / src
        if not p.exists():
            continue
        o = BUILD / (src + ".o")
        jobs.append([HIPCC, *HIP_FLAGS, "-c", str(p), "-o", str(o)])
        objs.append(o)
    for src in CXX_SOURCES:
        p = CSRC / src
        if not p.exists():
            continue
        o = BUILD / (src + ".o")
        jobs.append(["g++", *CXX_FLAGS, "-c", str(p), "-o", str(o)])
        objs.append(o)
    procs = [subprocess.Popen(j, stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True) for j in jobs]
    failed = []
    for j, p in zip(jobs, procs):
        out, errs = p.communicate()
        if verbose and (out or errs):
            sys.stderr.write(out + errs)
        if p.returncode != 0:
            failed.append((j, out + errs))
    if failed:
        for j, msg in failed:
            sys.stderr.write(" ".join(j) + "\n" + msg + "\n")
        raise RuntimeError("native build failed")
    tmp = LIB.with_suffix(".so.tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs), "-lz", "-lpthread"])
    os.replace(tmp, LIB)
    tree_src = CSRC / "kmc_tree_main.cpp"
    if tree_src.exists():
        _run(["g++", *CXX_FLAGS, str(tree_src), "-o", str(TREE), f"-L{PKG}", "-lkmc",
              f"-Wl,-rpath,$ORIGIN", "-lpthread"])
    return LIB


if __name__ == "__main__":
    build(verbose="-v" in sys.argv)
    print(LIB)
