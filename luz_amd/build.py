"""Build recipe for the native parts of luz_amd.

Compiles the HIP sources under luz_amd/csrc for gfx950 into
luz_amd/_lib/libluzrt.so (in-tree, so the built library travels with the
repository snapshot) and, when asked, the CPU oracle under oracle/_build.

    python -m luz_amd.build            # library only
    python -m luz_amd.build --oracle   # library + oracle
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
ROOT = PKG.parent
CSRC = PKG / "csrc"
LIBDIR = PKG / "_lib"
OBJDIR = ROOT / "_build" / "obj"
LIBNAME = "libluzrt.so"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("LUZ_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off keeps every FP32 multiply and add separately rounded (the
# visibility results are reproduced bit for bit on the CPU); packed FP32
# instructions are disabled for the device.
HIP_FLAGS = [
    "-std=c++17", "-O3", "-fPIC", f"--offload-arch={ARCH}",
    "-ffp-contract=off", "-fno-fast-math",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable", "-Wno-unused-value",
    "-Wno-unused-result",
    "-I", str(ROOT / "include"),
]
HOST_FLAGS = ["-std=c++17", "-O2", "-fPIC", "-ffp-contract=off", "-Wall", "-I", str(ROOT / "include")]


def _run(cmd: list[str]) -> None:
    proc = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in proc.stdout.splitlines() if "packed-fp32-ops" not in l)
    if proc.returncode != 0:
        raise RuntimeError(f"command failed ({proc.returncode}): {' '.join(cmd)}\n{out}")
    if out.strip():
        print(out, file=sys.stderr)


def build_library(jobs: int = 8) -> Path:
    OBJDIR.mkdir(parents=True, exist_ok=True)
    LIBDIR.mkdir(parents=True, exist_ok=True)
    hip_srcs = sorted(CSRC.glob("*.hip"))
    cpp_srcs = sorted(CSRC.glob("*.cpp"))
    jobs_list = []
    for src in hip_srcs:
        obj = OBJDIR / (src.stem + ".hip.o")
        jobs_list.append([HIPCC, "-x", "hip", *HIP_FLAGS, "-c", str(src), "-o", str(obj)])
    for src in cpp_srcs:
        obj = OBJDIR / (src.stem + ".cpp.o")
        jobs_list.append(["g++", *HOST_FLAGS, "-c", str(src), "-o", str(obj)])
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        list(ex.map(_run, jobs_list))
    objs = [j[-1] for j in jobs_list]
    tmp = LIBDIR / (LIBNAME + ".tmp")
    _run([HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *objs, "-lpthread"])
    out = LIBDIR / LIBNAME
    os.replace(tmp, out)
    return out


def build_oracle() -> Path:
    odir = ROOT / "oracle"
    _run(["make", "-C", str(odir), "-s"])
    return odir / "_build" / "libluz_oracle.so"


def main() -> None:
    ap = argparse.ArgumentParser()
    ap.add_argument("--oracle", action="store_true")
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args()
    print(build_library(a.j))
    if a.oracle:
        print(build_oracle())


if __name__ == "__main__":
    main()
