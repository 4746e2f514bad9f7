"""Build recipe of the native parts (no cmake/ninja needed).

Outputs (in-tree, git-ignored, shipped with the tree to the GPU machine):
  gpet_amd/libgpet.so    HIP kernels for gfx950 + host C++ + the C ABI
  gpet_amd/bin/gpet      drop-in command line (links libgpet.so)
  gpet_amd/data/tables/  cross-section tables in the reference's text format
  oracle/liboracle.so    CPU oracle (test infrastructure, built by oracle/Makefile)

Every HIP translation unit is compiled with -ffp-contract=off and correctly
rounded FP32 division/sqrt so that results match the host oracle bit for bit,
and with packed FP32 instructions disabled.
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
BUILD = os.path.join(HERE, "_build")
ARCH = os.environ.get("GPET_OFFLOAD_ARCH", "gfx950")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
if not os.path.exists(HIPCC):
    HIPCC = shutil.which("hipcc") or "hipcc"

COMMON = ["-O3", "-fPIC", "-std=c++17", "-ffp-contract=off", "-fno-fast-math", "-Wall",
          "-Wno-unused-function", "-I", os.path.join(ROOT, "include")]
HIP_FLAGS = COMMON + [
    "--offload-arch=" + ARCH,
    "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-munsafe-fp-atomics",
]

HIP_SOURCES = [
    os.path.join(CSRC, "host", "ctx.hip"),
    os.path.join(CSRC, "kernels", "transport.hip"),
    os.path.join(CSRC, "kernels", "digitizer.hip"),
    os.path.join(CSRC, "kernels", "mlp.hip"),
    os.path.join(CSRC, "host", "surrogate.hip"),
]
CPP_SOURCES = [os.path.join(CSRC, "host", "io.cpp")]


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if r.returncode != 0:
        raise RuntimeError("command failed (%d): %s\n%s" % (r.returncode, " ".join(cmd), r.stdout))
    return r.stdout


def _obj(src):
    return os.path.join(BUILD, os.path.basename(src) + ".o")


def build_library(verbose: bool = False) -> str:
    os.makedirs(BUILD, exist_ok=True)
    jobs = []
    for s in HIP_SOURCES:
        jobs.append([HIPCC, "-x", "hip"] + HIP_FLAGS + ["-c", s, "-o", _obj(s)])
    for s in CPP_SOURCES:
        jobs.append(["g++"] + COMMON + ["-c", s, "-o", _obj(s)])
    workers = min(len(jobs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(workers) as ex:
        for out in ex.map(_run, jobs):
            if verbose and out.strip():
                print(out)
    lib = os.path.join(HERE, "libgpet.so")
    tmp = lib + ".tmp"
    _run([HIPCC, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp]
         + [_obj(s) for s in HIP_SOURCES + CPP_SOURCES])
    os.replace(tmp, lib)
    bindir = os.path.join(HERE, "bin")
    os.makedirs(bindir, exist_ok=True)
    _run(["g++", "-O2", "-std=c++17", "-I", os.path.join(ROOT, "include"),
          os.path.join(CSRC, "cli", "gpet_main.cpp"), "-o", os.path.join(bindir, "gpet"),
          "-L", HERE, "-lgpet", "-Wl,-rpath,$ORIGIN/.."])
    return lib


def build_oracle() -> str:
    _run(["make", "-s", "-B", "-C", os.path.join(ROOT, "oracle")])
    return os.path.join(ROOT, "oracle", "liboracle.so")


def build_data() -> str:
    sys.path.insert(0, ROOT)
    from gpet_amd.data.prepare_data import materialize

    return materialize()


def build_all(verbose: bool = False) -> None:
    build_data()
    build_library(verbose)
    build_oracle()


if __name__ == "__main__":
    build_all(verbose="-v" in sys.argv)
    print("built", os.path.join(HERE, "libgpet.so"))
