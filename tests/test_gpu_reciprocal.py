"""The float64 reciprocal of the numba hit predicate (a hardware estimate and
two Newton steps, lem_device.h recip_f32_to_f64) equals the IEEE quotient
1.0 / (double)a for every finite non-zero float32 bit pattern (NaN for NaN):
F = 1.0 / A in rayfunctions.py:58-60 is reproduced bit for bit.  For an
infinite A (NaN there, 0 in IEEE) the predicate's decisions are the same,
checked on scenes whose products overflow float32."""
import numpy as np
import ctypes
import os

import pytest

pytestmark = pytest.mark.gpu

LIB = os.path.join(os.path.dirname(__file__), "native", "librecip_check.so")


def test_reciprocal_exact_for_every_float32():
    import torch

    assert os.path.exists(LIB), "tests/native/librecip_check.so missing: run the build first"
    lib = ctypes.CDLL(LIB)
    lib.lem_test_recip_check.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_void_p,
                                         ctypes.c_void_p, ctypes.c_void_p]
    lib.lem_test_recip_check.restype = ctypes.c_int
    dev = torch.device("cuda", 0)
    bad = torch.zeros(1, dtype=torch.int64, device=dev)
    first = torch.full((1,), -1, dtype=torch.int32, device=dev)
    stream = torch.cuda.current_stream(dev).cuda_stream
    chunk = 1 << 28
    for c in range(16):
        rc = lib.lem_test_recip_check(c * chunk, chunk, bad.data_ptr(), first.data_ptr(), stream)
        assert rc == 0
        torch.cuda.synchronize(dev)
        print(f"reciprocal check: patterns up to {(c + 1) * chunk:#x}, mismatches {int(bad.item())}",
              flush=True)
    assert int(bad.item()) == 0, f"first mismatching float32 bits {int(first.item()) & 0xffffffff:#010x}"


def test_overflowing_products_match_oracle(native, oracle):
    """Triangles with edges near 5e19 make the float32 determinant A overflow
    to inf; rays and smaller triangles of about 1e12 (no overflow in their own
    products) give real hits.  Masks and distances equal the oracle's (IEEE
    1/A = 0 there)."""
    rng = np.random.default_rng(3)
    L = 1e12
    big = (rng.uniform(-L, L, (100, 1, 3)) + rng.uniform(-5e19, 5e19, (100, 3, 3)))
    small = (rng.uniform(-L, L, (60, 1, 3)) + rng.uniform(-0.4 * L, 0.4 * L, (60, 3, 3)))
    tris = np.concatenate([big, small]).reshape(-1, 9).astype(np.float32)
    src = rng.uniform(-L, L, (40, 3)).astype(np.float32)
    snk = rng.uniform(-L, L, (70, 3)).astype(np.float32)
    e1 = tris[:100, 3:6].astype(np.float64) - tris[:100, 0:3]
    e2 = tris[:100, 6:9].astype(np.float64) - tris[:100, 0:3]
    assert (np.linalg.norm(e1, axis=1) * np.linalg.norm(e2, axis=1) > 3.5e38).mean() > 0.5
    bits, dist = native.visibility_dense(src, snk, tris, any_hit=False, want_dist=True)
    hit, odist = oracle.visibility_dense(src, snk, tris)
    assert 0.05 < hit.mean() < 0.95
    assert np.array_equal(native.unpack_bits(bits, len(snk)), hit)
    assert np.array_equal(dist.cpu().numpy(), odist)
    g = native.Grid(tris)
    bits_g, dist_g = g.visibility(src, snk, any_hit=False, want_dist=True)
    assert np.array_equal(native.unpack_bits(bits_g, len(snk)), hit)
    assert np.array_equal(dist_g.cpu().numpy(), odist)
