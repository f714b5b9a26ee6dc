"""Builds libmgx.so (HIP for gfx950) and the mm_1d / mm_2d executables in-tree.

Usage: python -m mgxbench_amd.build [--jobs N]

Every HIP source is compiled by hipcc for --offload-arch=gfx950 with packed
FP32 instructions disabled; objects go to mgxbench_amd/_build/, the shared
library to mgxbench_amd/libmgx.so and the executables to mgxbench_amd/bin/.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
BUILD = PKG / "_build"
BIN = PKG / "bin"
LIB = PKG / "libmgx.so"
INCLUDE = PKG.parent / "include"

LIB_SOURCES = ["runtime.hip", "dgemm.hip", "ozaki.hip", "summa.hip", "accuracy.hip", "mlp.hip", "tune.hip",
               "ceiling.hip", "noalloc.hip"]
EXE_SOURCES = {"mm_1d": "mm_1d.cpp", "mm_2d": "mm_2d.cpp"}  # linked against libmgx.so

ARCH = "gfx950"
COMMON_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    f"-I{INCLUDE}",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and Path(cand).exists():
            return cand
    raise RuntimeError("hipcc not found: libmgx needs the ROCm HIP compiler")


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    # The host pass does not know the packed-fp32 feature; drop that one warning.
    out = "\n".join(l for l in proc.stdout.splitlines()
                    if "packed-fp32-ops' is not a recognized feature" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build(jobs: int = 8) -> Path:
    cc = hipcc()
    BUILD.mkdir(exist_ok=True)
    BIN.mkdir(exist_ok=True)
    objs = []
    cmds = []
    for src in LIB_SOURCES:
        obj = BUILD / (Path(src).stem + ".o")
        objs.append(obj)
        cmds.append([cc, *COMMON_FLAGS, "-c", str(CSRC / src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for f in [ex.submit(_run, c) for c in cmds]:
            f.result()
    tmp = LIB.with_suffix(".so.tmp")
    _run([cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, objs)])
    os.replace(tmp, LIB)
    exe_cmds = []
    for name, src in EXE_SOURCES.items():
        exe = BIN / name
        exe_cmds.append([cc, "-x", "hip", *COMMON_FLAGS, str(CSRC / src), "-o", str(exe),
                         f"-L{PKG}", "-lmgx", "-Wl,-rpath,$ORIGIN/.."])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        for f in [ex.submit(_run, c) for c in exe_cmds]:
            f.result()
    return LIB


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--jobs", type=int, default=8)
    args = ap.parse_args()
    print(build(args.jobs))


if __name__ == "__main__":
    main()
