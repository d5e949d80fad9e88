"""Build liblem.so (the gfx950 HIP library) in-tree.

    python lyceanem-python_amd/build_lib.py [--arch gfx950] [-j N]

Each csrc/*.hip is compiled to an object with hipcc and the objects are linked
into lyceanem/_lib/liblem.so, which lyceanem/_native.py loads with ctypes.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT_DIR = os.path.join(HERE, "lyceanem", "_lib")
BUILD_DIR = os.path.join(HERE, "_build")
LIB = os.path.join(OUT_DIR, "liblem.so")

# -ffp-contract=off + correctly rounded fp32 div/sqrt: the float32 predicate
# arithmetic must match the reference expression by expression.
# -packed-fp32-ops: packed float32 VALU instructions stay disabled on this pool.
COMMON = [
    "-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function",
    "-ffp-contract=off", "-fhip-fp32-correctly-rounded-divide-sqrt",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
]


def hipcc() -> str:
    path = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(path):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return path


def compile_one(src: str, arch: str) -> str:
    obj = os.path.join(BUILD_DIR, os.path.basename(src) + ".o")
    cmd = [hipcc(), f"--offload-arch={arch}", *COMMON, "-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def build(arch: str = "gfx950", jobs: int = 8, verbose: bool = False) -> str:
    os.makedirs(BUILD_DIR, exist_ok=True)
    os.makedirs(OUT_DIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(lambda s: compile_one(s, arch), srcs))
    tmp = LIB + ".tmp"
    cmd = [hipcc(), f"--offload-arch={arch}", "-shared", "-fPIC", *objs, "-o", tmp]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} from {len(srcs)} sources for {arch}")
    return LIB


TEST_NATIVE = os.path.join(os.path.dirname(HERE), "tests", "native")


def build_test_libs(arch: str = "gfx950") -> list:
    """Check libraries of the GPU tests: tests/native/<name>.hip ->
    tests/native/lib<name>.so (not part of the package)."""
    out = []
    for src in sorted(glob.glob(os.path.join(TEST_NATIVE, "*.hip"))):
        lib = os.path.join(TEST_NATIVE, "lib" + os.path.basename(src)[:-4] + ".so")
        cmd = [hipcc(), f"--offload-arch={arch}", *COMMON, "-shared", src, "-o", lib + ".tmp"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
        os.replace(lib + ".tmp", lib)
        out.append(lib)
    return out


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--arch", default=os.environ.get("LEM_ARCH", "gfx950"))
    ap.add_argument("-j", type=int, default=8)
    a = ap.parse_args(argv)
    build(a.arch, a.j, verbose=True)
    build_test_libs(a.arch)
    return 0


if __name__ == "__main__":
    sys.exit(main())
