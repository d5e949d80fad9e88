"""CPU build of the grid traversal (lyceanem-python_amd/csrc/lem_grid_walk.h) and
of the grid build's per-item arithmetic, checked ray by ray against the
oracle's brute force over every triangle -- no GPU needed.

tests/native/grid_walk_host.cpp builds grids by host loops from the functions
the build kernels call and casts every ray the way each traversal kernel casts
it (float macro boxes, the quantised LDS table for ordinary and plane-parallel
rays, the walk from the sink end, both predicates), on terrains whose rays lie
in vertex planes (shared-edge near-ties that the kernel1Dv3 distance replay
must resolve), points exactly on the grid's float32 cell boundaries, rays
along grid planes and zero-length rays, origins 300 m to 30 km from a 60 m
scene, a terrain under a tall empty sky (long runs of empty macro blocks),
macro blocks of 1 to 16 cells and 1 to 4 cells per triangle.  A second
build runs under AddressSanitizer and UndefinedBehaviorSanitizer (including
float-to-int conversion overflow)."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
NATIVE = os.path.join(ROOT, "tests", "native")
ORACLE_C = os.path.join(ROOT, "oracle", "lem_oracle.c")
SRC = os.path.join(NATIVE, "grid_walk_host.cpp")
FP = ["-ffp-contract=off", "-fno-fast-math"]


def _build(tmp, name, opt, extra):
    gcc, gxx = shutil.which("gcc"), shutil.which("g++")
    if not gcc or not gxx:
        pytest.skip("gcc / g++ not available")
    orc = os.path.join(tmp, name + "_oracle.o")
    exe = os.path.join(tmp, name)
    subprocess.run([gcc, opt, *FP, "-std=c11", *extra, "-c", ORACLE_C, "-o", orc], check=True,
                   capture_output=True)
    r = subprocess.run([gxx, opt, "-std=c++17", *FP, "-Wall", "-Wno-unknown-pragmas", *extra,
                        "-I", NATIVE, SRC, orc, "-o", exe, "-lm"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe, *args):
    r = subprocess.run([exe, *args], capture_output=True, text=True, timeout=600)
    lines = r.stdout.strip().splitlines()
    bad = [ln for ln in lines if " 0 mask and 0 distance" not in ln and "total" not in ln]
    print("\n".join(lines[-1:] + bad[:20]))
    return r, lines, bad


def test_grid_walk_host_matches_oracle(tmp_path):
    exe = _build(str(tmp_path), "walk", "-O2", ["-fopenmp"])
    r, lines, bad = _run(exe)
    assert r.returncode == 0 and not bad, r.stderr
    assert lines[-1].startswith("total:") and lines[-1].endswith(" 0 mismatches")
    rays = int(lines[-1].split()[1])
    assert rays > 900_000


def test_grid_walk_host_sanitizers(tmp_path):
    san = ["-g", "-fno-omit-frame-pointer",
           "-fsanitize=address,undefined,float-cast-overflow", "-fno-sanitize-recover=all"]
    exe = _build(str(tmp_path), "walk_san", "-O1", san)
    r, lines, bad = _run(exe, "quick")
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr[-3000:]
    assert r.returncode == 0 and not bad
    assert lines[-1].endswith(" 0 mismatches")
