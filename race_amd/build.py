"""Build librace_amd.so (HIP for gfx950) in-tree with hipcc.

    python -m race_amd.build            # build the library
    python -m race_amd.build --check    # build and list exported symbols
"""
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_obj")
LIB = os.path.join(HERE, "librace_amd.so")
ARCH = os.environ.get("RACE_AMD_ARCH", "gfx950")

SOURCES = ["engine.cpp", "bundle.cpp", "simdify.cpp", "bfs.hip", "exec.hip", "bundle_exec.hip", "sell.hip", "refine.hip", "mlp.hip", "schedule_io.cpp", "rcm.cpp", "features.cpp", "capi.cpp"]

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"

COMMON = [
    "-O3", "-fPIC", "-std=c++17", f"--offload-arch={ARCH}",
    # packed FP32 instructions stay off (they are generated from plain float code otherwise)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
]


# host callbacks run on OpenMP threads (capi.cpp), as in the reference
OPENMP_SOURCES = {"capi.cpp"}


def _compile(src):
    out = os.path.join(OBJ, os.path.splitext(src)[0] + ".o")
    flags = COMMON + (["-fopenmp"] if src in OPENMP_SOURCES else [])
    cmd = [HIPCC] + flags + ["-c", os.path.join(CSRC, src), "-o", out]
    if src.endswith(".cpp"):
        cmd = [HIPCC] + flags + ["-x", "hip", "-c", os.path.join(CSRC, src), "-o", out]
    subprocess.check_call(cmd)
    return out


def build(verbose=False):
    os.makedirs(OBJ, exist_ok=True)
    with ThreadPoolExecutor(max_workers=min(4, len(SOURCES))) as ex:
        objs = list(ex.map(_compile, SOURCES))
    tmp = LIB + ".tmp"
    subprocess.check_call([HIPCC, "-shared", "-fPIC", "-fopenmp", f"--offload-arch={ARCH}", "-o", tmp] + objs)
    os.replace(tmp, LIB)
    if verbose:
        print("built", LIB)
    build_drop_in()
    return LIB


DROP_IN_SRC = os.path.join(os.path.dirname(HERE), "tests", "cpp", "drop_in.cpp")
DROP_IN_BIN = os.path.join(os.path.dirname(HERE), "tests", "cpp", "drop_in")


def build_drop_in():
    """C++ program using RACE::Interface (include/race_amd_interface.hpp)."""
    inc = os.path.join(os.path.dirname(HERE), "include")
    subprocess.check_call([HIPCC, "-std=c++17", "-O2", f"--offload-arch={ARCH}", "-I", inc, DROP_IN_SRC,
                           "-L", HERE, "-lrace_amd", "-Wl,-rpath,$ORIGIN/../../race_amd",
                           "-o", DROP_IN_BIN])


if __name__ == "__main__":
    build(verbose=True)
    if "--check" in sys.argv:
        out = subprocess.check_output(["nm", "-D", "--defined-only", LIB]).decode()
        print("\n".join(l for l in out.splitlines() if " race_" in l))
