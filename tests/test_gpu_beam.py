"""GPU test of MaximumDirectivityMap (MFMA beam-power kernel) against the
oracle's per-command loop (beamforming.py:484-673)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _array_patterns(rng, n_side, az, el, wavelength):
    """Element patterns of an n x n array at lambda/2 in z = 0: an element
    factor (cos of the angle off +z, clipped) times the array phase, plus a
    little random coupling."""
    g = (np.arange(n_side) - (n_side - 1) / 2) * wavelength / 2
    X, Y = np.meshgrid(g, g)
    pts = np.stack([X.ravel(), Y.ravel(), np.zeros(n_side * n_side)], 1)
    A, E = np.meshgrid(np.radians(az), np.radians(el))
    u = np.stack([np.cos(E) * np.cos(A), np.cos(E) * np.sin(A), np.sin(E)], -1)
    k = 2 * np.pi / wavelength
    phase = np.exp(1j * k * np.einsum("nc,ijc->nij", pts, u))
    ef = np.clip(u[..., 2], 0.05, None)
    n = len(pts)
    noise = lambda: 0.05 * (rng.normal(size=(n,) + A.shape) + 1j * rng.normal(size=(n,) + A.shape))
    Et = (ef * phase + noise()).astype(np.complex64)
    Ep = (0.5 * np.cos(A) * phase + noise()).astype(np.complex64)
    return pts, Et, Ep


@pytest.mark.parametrize("forming,tsa", [("Total", 4 * np.pi), ("Etheta", 2 * np.pi)])
def test_max_directivity_map_matches_oracle(native, oracle, forming, tsa):
    from lyceanem.electromagnetics.beamforming import MaximumDirectivityMap

    print(f"max directivity {forming}", flush=True)
    rng = np.random.default_rng(4)
    az, el = np.linspace(-180.0, 180.0, 25), np.linspace(-90.0, 90.0, 13)
    pts, Et, Ep = _array_patterns(rng, 4, az, el, 0.1)
    got = MaximumDirectivityMap(Et, Ep, pts, 0.1, az_range=az, elev_range=el, forming=forming,
                                total_solid_angle=tsa)
    ref = oracle.max_directivity_map(Et, Ep, pts, 0.1, az, el, forming=forming,
                                     total_solid_angle=tsa)
    assert got.shape == ref.shape == (13, 25, 3)
    np.testing.assert_allclose(got, ref, rtol=2e-4, atol=2e-4 * np.abs(ref).max())
    assert ref[..., 2].max() > 5  # a real array gain


def test_beam_power_rejects_bad_shapes(native):
    with pytest.raises(native.LemError):
        native.beam_power(np.ones((3, 4)), np.ones((5, 7)), np.ones((5, 7)), np.ones(7),
                          np.zeros(3))
