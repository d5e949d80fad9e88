"""GPU test of lem_field_los_noalloc, the allocation-free variant of
lem_field_los (its source-sum partials live in a static device array instead
of a caller workspace): called through the C-ABI several times in a row on
shrinking inputs, each result against lem_field_los on the same inputs."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _inputs(native, S, K, seed):
    import torch

    rng = np.random.default_rng(seed)
    dev = native.device()
    src = torch.from_numpy(rng.uniform(-1, 1, (S, 3)).astype(np.float32)).to(dev)
    snk = torch.from_numpy((rng.uniform(-1, 1, (K, 3)) * 30 + [0, 0, 60]).astype(np.float32)).to(dev)
    nrm = torch.from_numpy(rng.normal(size=(K, 3)).astype(np.float32)).to(dev)
    w = torch.from_numpy((rng.normal(size=(S, 3)) + 1j * rng.normal(size=(S, 3)))
                         .astype(np.complex64)).to(dev)
    wk = torch.from_numpy((rng.normal(size=(K, 3)) + 1j * rng.normal(size=(K, 3)))
                          .astype(np.complex64)).to(dev)
    g = torch.Generator(device="cpu").manual_seed(seed)
    bits = torch.randint(-2 ** 62, 2 ** 62, (S, (K + 63) // 64), generator=g,
                         dtype=torch.int64).to(dev)
    bits_sm = torch.randint(-2 ** 62, 2 ** 62, (K, (S + 63) // 64), generator=g,
                            dtype=torch.int64).to(dev)
    return src, snk, nrm, w, wk, bits, bits_sm


def _noalloc(native, src, w, snk, nrm, wk, bits, alpha, beta, flags):
    import torch

    lib = native.load_library()
    S, K = src.shape[0], snk.shape[0]
    out = torch.full((K, 3), float("nan"), dtype=torch.complex64, device=src.device)
    p = native.ptr
    native.check(lib.lem_field_los_noalloc(p(src), p(w), S, p(snk), p(nrm), p(wk), K, p(bits),
                                           float(alpha), float(beta), flags, p(out),
                                           native.stream_ptr()), "lem_field_los_noalloc")
    return out.cpu().numpy()


def _compare(got, ref, exact, bound):
    """Bit-identical when the chunking is lem_field_los's own; regrouped, the
    per-sink difference is float32 summation rounding: below 1e-5 of (a bound
    of) the sum of the sink's |terms|."""
    assert np.isfinite(got).all()
    if exact:
        assert np.array_equal(got, ref)
        return
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    per = (np.abs(got - ref).max(axis=-1) / bound).max()
    print(f"  regrouped partial sums: normwise {err:.2e}, per-sink {per:.2e}", flush=True)
    assert err < 1e-5 and per < 1e-5


def test_noalloc_equals_field_los(native, term_bound):
    lib = native.load_library()
    kmax = lib.lem_field_los_noalloc_max_sinks()
    # (S, K): the first call is too wide for lem_field_los's chunk count to fit
    # the scratch (fewer chunks, sums regrouped); the next ones are smaller and
    # use lem_field_los's own chunking (bit-identical)
    for S, K, seed in [(4096, 65341, 1), (3000, 5000, 2), (700, 333, 3)]:
        src, snk, nrm, w, wk, bits, bits_sm = _inputs(native, S, K, seed)
        nc = lib.lem_field_los_workspace(S, K, 0) // (2 * 6 * K * 4)
        exact = 12 * K * nc * 4 <= (16 << 20)
        assert K <= kmax
        for flags, alpha, use_wk, sm in [(0, 0.0, False, False),
                                         (native.FIELD_SINK_TANGENT, 0.3, True, False),
                                         (native.FIELD_BITS_SINK_MAJOR, 0.0, True, True)]:
            print(f"noalloc S={S} K={K} flags={flags} alpha={alpha} chunks={nc} "
                  f"exact={exact}", flush=True)
            b = bits_sm if sm else bits
            ref = native.field_los(src, w, snk, nrm, hit_bits=b, snk_w=wk if use_wk else None,
                                   alpha=alpha, beta=2 * np.pi / 0.2,
                                   sink_tangent=bool(flags & native.FIELD_SINK_TANGENT),
                                   bits_sink_major=sm).cpu().numpy()
            got = _noalloc(native, src, w, snk, nrm, wk if use_wk else None, b, alpha,
                           2 * np.pi / 0.2, flags)
            _compare(got, ref, exact, None if exact else term_bound(
                src, w, snk, nrm, wk if use_wk else None, alpha, 2 * np.pi / 0.2))
    # the per-element network and an empty source set pass through
    src, snk, nrm, w, wk, bits, _ = _inputs(native, 64, 100, 4)
    got = _noalloc(native, src[:0], w[:0], snk, nrm, None, None, 0.0, 1.0, 0)
    assert np.array_equal(got, np.zeros_like(got))
    # sizes beyond the scratch are rejected with a message
    src, snk, nrm, w, wk, bits, _ = _inputs(native, 8, kmax + 1, 5)
    with pytest.raises(native.LemError, match="scratch"):
        _noalloc(native, src, w, snk, nrm, None, bits, 0.0, 1.0, 0)
