"""CPU checks of the beamforming oracle and the host parts of
electromagnetics.beamforming: closed-form directivity of an isotropic pattern
under the reference's power sum, the first scheme's 4 pi solid angle, and the
vectorised power weights against the literal cell loop."""
import numpy as np
import pytest

AZ = np.linspace(-180.0, 180.0, 19)
EL = np.linspace(-90.0, 90.0, 19)


def _isotropic_directivity(az, el, tsa=4 * np.pi):
    """U / Uav of a constant pattern: power_sum = sum over cells of
    |sin(radians(az_j + az_j+1) / 2)| (unit corner powers)."""
    s = np.abs(np.sin(np.radians(az[:-1] + az[1:]) / 2.0)).sum() * (len(el) - 1)
    return tsa / (s * np.radians(az[1] - az[0]) * np.radians(el[1] - el[0]))


def test_isotropic_directivity_closed_form(oracle):
    E = np.ones((len(EL), len(AZ)), np.complex64)
    dt, dp, dtot = oracle.directivity_transform(E, 0 * E, AZ, EL)
    np.testing.assert_allclose(dt, _isotropic_directivity(AZ, EL), rtol=1e-12)
    assert not dp.any()
    np.testing.assert_allclose(dtot, dt, rtol=1e-12)


@pytest.mark.parametrize("tsa,factor", [(2 * np.pi, 1.0), (8 * np.pi, 2.0)])
def test_first_scheme_uses_four_pi(oracle, tsa, factor):
    """One isotropic element: all three schemes give the same pattern; the
    equal-gain Etheta scheme is normalised by 4 pi, the others by
    total_solid_angle, and the larger directivity is kept."""
    E = np.exp(1j * 0.3) * np.ones((1, len(EL), len(AZ)), np.complex64)
    m = oracle.max_directivity_map(E, 0 * E, np.zeros((1, 3)), 0.1, AZ, EL,
                                   total_solid_angle=tsa)
    np.testing.assert_allclose(m[..., 0], factor * _isotropic_directivity(AZ, EL), rtol=1e-6)
    np.testing.assert_allclose(m[..., 2], m[..., 0], rtol=1e-12)


def test_power_weights_match_cell_loop(oracle):
    from lyceanem.electromagnetics.beamforming import _power_weights, directivity_transform

    rng = np.random.default_rng(1)
    az, el = np.linspace(-170, 170, 23), np.linspace(-80, 85, 17)
    Et = rng.normal(size=(17, 23)) + 1j * rng.normal(size=(17, 23))
    Ep = rng.normal(size=(17, 23)) + 1j * rng.normal(size=(17, 23))
    ref = oracle.directivity_transform(Et, Ep, az, el, 3.0)
    got = directivity_transform(Et, Ep, az, el, 3.0)
    for a, b in zip(ref, got[:3]):
        np.testing.assert_allclose(b, a, rtol=1e-6)
    q = _power_weights(az, el)
    assert q.shape == (17 * 23,)
    s = np.abs(np.sin(np.radians(az[:-1] + az[1:]) / 2.0))
    assert q.sum() == pytest.approx(s.sum() * 16, rel=1e-12)


def test_weight_builders():
    from lyceanem.electromagnetics.beamforming import WavefrontWeights, WeightTruncation

    # broadside steering of a planar array in z = 0: equal phases
    pts = np.stack([np.arange(8) * 0.05, np.zeros(8), np.zeros(8)], 1)
    w = WavefrontWeights(pts, np.array([[0, 0, 1.0]]), 0.1)
    np.testing.assert_allclose(np.angle(w), 0.0, atol=1e-4)
    # endfire along +x: phase grows by k d per element
    w = WavefrontWeights(pts, np.array([[1.0, 0, 0]]), 0.1)
    dphi = np.angle(w[1:] / w[:-1])
    np.testing.assert_allclose(np.abs(dphi), np.pi, atol=1e-3)
    # 1-bit truncation maps phases onto {-pi, pi}
    t = WeightTruncation(np.exp(1j * np.array([0.1, 2.0, -2.0])), 1)
    np.testing.assert_allclose(np.abs(np.imag(t)), 0.0, atol=1e-12)


def test_command_weights_match_per_command_builders():
    from lyceanem.electromagnetics import beamforming as BF

    rng = np.random.default_rng(2)
    az, el = np.linspace(-180, 180, 13), np.linspace(-90, 90, 7)
    Et = (rng.normal(size=(5, 7, 13)) + 1j * rng.normal(size=(5, 7, 13))).astype(np.complex64)
    Ep = (rng.normal(size=(5, 7, 13)) + 1j * rng.normal(size=(5, 7, 13))).astype(np.complex64)
    pts = rng.uniform(-0.2, 0.2, (5, 3))
    ai, ei = np.arange(0, 13, 2), np.arange(7)
    W, pstar = BF.command_weights(Et, Ep, pts, 0.1, az, el, ai, ei)
    for a, ia in enumerate(ai):
        for e, ie in enumerate(ei):
            cmd = np.array([az[ia], el[ie]], np.float32)
            sv = np.array([BF.sph2cart(np.radians(cmd[0]), np.radians(cmd[1]), 1)])
            np.testing.assert_allclose(W[a, e, 0], BF.EGCWeights(Et, Ep, cmd, "Etheta", az, el),
                                       atol=1e-6)
            np.testing.assert_allclose(W[a, e, 1], BF.EGCWeights(Et, Ep, cmd, "Ephi", az, el),
                                       atol=1e-6)
            np.testing.assert_allclose(W[a, e, 2], BF.WavefrontWeights(pts, sv, 0.1), atol=1e-6)
            assert (pstar[a, e] == e * 13 + a).all()


def test_gpu_beamforming_map_launch_form():
    """The reference's GPUBeamformingMap is an unfinished kernel that zeroes
    the map inside (az_range, el_range); the launch form [blocks, threads] and
    that result are kept."""
    from lyceanem.electromagnetics.beamforming import GPUBeamformingMap

    m = np.ones((5, 4))
    GPUBeamformingMap[(1, 1), (16, 16)](None, None, m, np.zeros(3), np.zeros(2), 0.1)
    assert np.all(m[:3, :2] == 0) and m[3:].all() and m[:, 2:].all()
