"""GPU tests of the batched-excitation MFMA contraction (lem_gw_mfma)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _scene():
    from lyceanem import scenes

    return scenes.batched_c5(n_src_side=12, n_el=8, n_az=12, n_freq=3)


def test_gw_mfma_vs_fp64(native, oracle):
    sc = _scene()
    src, snk, snk_n = sc["src"], sc["snk"], sc["snk_n"]
    S, K = len(src), len(snk)
    pad = (-S) % 32
    src_p = np.concatenate([src, np.tile(src[:1], (pad, 1))])
    W = np.concatenate([sc["W"], np.zeros((sc["W"].shape[0], pad, 64), np.complex64)], 1)
    E = native.gw_mfma(src_p, sc["pol"], snk, snk_n, sc["wavenumbers"], W).cpu().numpy()
    assert E.shape == (3, 3, K, 64)
    worst = 0.0
    for f, kw in enumerate(sc["wavenumbers"]):
        G = oracle.field_los(src, sc["pol"].astype(np.complex64), snk, snk_n, 0.0, kw,
                             elements=True)  # (S, K, 3) float64
        ref = np.einsum("skc,sb->ckb", G, sc["W"][f].astype(np.complex128))
        err = np.linalg.norm(E[f] - ref) / np.linalg.norm(ref)
        worst = max(worst, err)
    print(f"G.W MFMA vs float64: worst normwise rel err {worst:.2e}", flush=True)
    assert worst < 1e-4


def test_gw_mfma_equals_per_excitation_kernel(native):
    sc = _scene()
    src, snk, snk_n = sc["src"][:128], sc["snk"], sc["snk_n"]
    W = sc["W"][:, :128]
    E = native.gw_mfma(src, sc["pol"], snk, snk_n, sc["wavenumbers"], W).cpu().numpy()
    for f in (0, 2):
        for b in (0, 17, 63):
            w = (W[f, :, b][:, None] * sc["pol"][None, :]).astype(np.complex64)
            e1 = native.field_los(src, w, snk, snk_n, beta=sc["wavenumbers"][f]).cpu().numpy()
            got = E[f, :, :, b].T
            assert np.linalg.norm(got - e1) / np.linalg.norm(e1) < 1e-5


def test_gw_mfma_rejects_bad_shapes(native):
    sc = _scene()
    with pytest.raises(native.LemError):
        native.gw_mfma(sc["src"][:100], sc["pol"], sc["snk"], sc["snk_n"], sc["wavenumbers"],
                       sc["W"][:, :100])
