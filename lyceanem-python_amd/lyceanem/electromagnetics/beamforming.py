"""Beamforming consumers of the element-pattern network
(lyceanem/electromagnetics/beamforming.py).

MaximumDirectivityMap keeps the reference signature and result; its hot loop
-- beamforming every element pattern with three weight sets per command angle
and integrating the resulting power over the whole pattern -- runs as one
FP32-MFMA contraction with a fused power epilogue (lem_beam.hip) instead of a
numpy loop over command angles.  The weight builders and directivity_transform
are host numpy restatements with the reference's arithmetic."""
from __future__ import annotations

import numpy as np

from .. import _native

LIGHT_SPEED = 299792458.0  # scipy.constants.speed_of_light


def sph2cart(az, el, r):
    """math_functions.py:46-52 (radians)."""
    rc = r * np.cos(el)
    return rc * np.cos(az), rc * np.sin(az), r * np.sin(el)


def WavefrontWeights(source_coords, steering_vector, wavelength):
    """Wavefront steering weights (beamforming.py:81-120): phase of the path
    difference towards a point 1e9 m along the steering vector."""
    p = np.asarray(source_coords, np.float64).reshape(-1, 3)
    s = np.asarray(steering_vector, np.float64).ravel() * 1e9
    dist = np.sqrt(np.abs((p[:, 0] - s[0]) ** 2 + (p[:, 1] - s[1]) ** 2 + (p[:, 2] - s[2]) ** 2))
    dist = dist - np.min(dist)
    delays = dist / LIGHT_SPEED
    return np.exp(1j * 2 * np.pi * (LIGHT_SPEED / wavelength) * delays).astype(np.complex64)


def EGCWeights(Etheta, Ephi, command_angles, polarization_switch="Etheta", az_range=None,
               elev_range=None):
    """Equal-gain-combining weights (beamforming.py:278-326): the conjugate phase
    of each element's pattern at the grid point nearest the command angle."""
    if az_range is None:
        az_range = np.linspace(-180.0, 180.0, 19)
    if elev_range is None:
        elev_range = np.linspace(-90.0, 90.0, 19)
    ai = np.argmin(np.abs(az_range - command_angles[0]))
    ei = np.argmin(np.abs(elev_range - command_angles[1]))
    E = Etheta if polarization_switch == "Etheta" else Ephi
    ang = -np.angle(np.asarray(E)[:, ei, ai].astype(np.complex64))
    return np.exp(1j * ang).astype(np.complex64)


def WeightTruncation(weights, resolution):
    """Phase quantisation to `resolution` bits (beamforming.py:1133-1147)."""
    quant = 2 ** resolution
    levels = np.round((quant - 1) * ((np.angle(weights) + np.pi) / (2 * np.pi))) / (quant - 1)
    return np.abs(weights) * np.exp(1j * ((levels * 2 * np.pi) - np.pi))


def _power_weights(az_range, elev_range):
    """Per-point weights q of directivity_transform's power sum
    (beamforming.py:1087-1110): each grid cell adds the mean of its four corner
    powers times |sin(radians(az_j + az_j+1) / 2)| (the reference's factor, the
    mean azimuth), so a point collects a quarter of that factor from each cell
    it is a corner of.  Returned el-major, az-fastest."""
    az = np.asarray(az_range, np.float64)
    n_el, n_az = len(elev_range), len(az)
    s = np.abs(np.sin(np.radians(az[:-1] + az[1:]) / 2.0))
    q = np.zeros((n_el, n_az))
    for di in (0, 1):
        for dj in (0, 1):
            q[di:n_el - 1 + di, dj:n_az - 1 + dj] += s[None, :] / 4.0
    return q.ravel()


def directivity_transform(Etheta, Ephi, az_range=None, elev_range=None,
                          total_solid_angle=(4 * np.pi)):
    """Directivity of a (n_el, n_az) pattern (beamforming.py:1037-1129):
    returns (Dtheta, Dphi, Dtot, Dmax)."""
    if az_range is None:
        az_range = np.linspace(-180.0, 180.0, 19)
    if elev_range is None:
        elev_range = np.linspace(-90.0, 90.0, 19)
    Ut = np.abs(np.asarray(Etheta) ** 2)
    Up = np.abs(np.asarray(Ephi) ** 2)
    Utot = Ut + Up
    q = _power_weights(az_range, elev_range).reshape(Ut.shape)
    power_sum = float(np.sum((Ut.astype(np.float32) + Up.astype(np.float32)) * q))
    Uav = power_sum * (np.radians(az_range[1] - az_range[0]) *
                       np.radians(elev_range[1] - elev_range[0])) / total_solid_angle
    Umax = np.array([np.nanmax(Ut), np.nanmax(Up), np.nanmax(Utot)], np.float32)
    return Ut / Uav, Up / Uav, Utot / Uav, Umax / Uav


def command_weights(Et, Ep, points, wavelength, az_range, elev_range, az_index, elev_index,
                    phase_resolution=24):
    """Weight rows of MaximumDirectivityMap, (n_az_index, n_elev_index, 3, ne)
    complex64 -- equal-gain combining on Etheta, on Ephi and wavefront steering
    for each command angle (beamforming.py:560-600) -- and the grid point each
    row's directivity is read at, (elev_inc, az_inc) flattened el-major."""
    ne, n_el, n_az = Et.shape
    cmd_az = np.asarray(az_range)[np.asarray(az_index)].astype(np.float32)
    cmd_el = np.asarray(elev_range)[np.asarray(elev_index)].astype(np.float32)
    ai = np.argmin(np.abs(np.asarray(az_range)[None, :] - cmd_az[:, None]), axis=1)
    ei = np.argmin(np.abs(np.asarray(elev_range)[None, :] - cmd_el[:, None]), axis=1)
    A, E = len(cmd_az), len(cmd_el)
    W = np.empty((A, E, 3, ne), np.complex64)
    # Et[:, ei, ai] for every (command az, command el)
    W[:, :, 0] = np.exp(1j * -np.angle(Et[:, ei[None, :], ai[:, None]])).transpose(1, 2, 0)
    W[:, :, 1] = np.exp(1j * -np.angle(Ep[:, ei[None, :], ai[:, None]])).transpose(1, 2, 0)
    # the reference's command angles are float32, so is its steering vector
    az_r = np.radians(cmd_az)[:, None]
    el_r = np.radians(cmd_el)[None, :]
    sv = np.stack(np.broadcast_arrays(*sph2cart(az_r, el_r, np.float32(1))), -1)
    sv = sv.reshape(-1, 3).astype(np.float64)
    p = np.asarray(points, np.float64).reshape(-1, 3)
    for i in range(0, len(sv), 4096):
        s = sv[i:i + 4096] * 1e9
        dist = np.sqrt(np.abs((p[None, :, 0] - s[:, None, 0]) ** 2 +
                              (p[None, :, 1] - s[:, None, 1]) ** 2 +
                              (p[None, :, 2] - s[:, None, 2]) ** 2))
        dist = dist - dist.min(axis=1, keepdims=True)
        w = np.exp(1j * 2 * np.pi * (LIGHT_SPEED / wavelength) * (dist / LIGHT_SPEED))
        W.reshape(-1, 3, ne)[i:i + 4096, 2] = w.astype(np.complex64)
    if phase_resolution <= 12:
        W = WeightTruncation(W, phase_resolution).astype(np.complex64)
    a_idx, e_idx = np.meshgrid(np.arange(A), np.arange(E), indexing="ij")
    pstar = np.repeat((e_idx * n_az + a_idx)[..., None], 3, axis=2).astype(np.int32)
    return W, pstar


class _ZeroMapKernel:
    """GPUBeamformingMap (beamforming.py:1553-1562) is unfinished in the
    reference: launched as GPUBeamformingMap[blocks, threads](...), every
    thread inside (az_range, el_range) writes DirectivityMap[az, el] = 0.  The
    same launch form and result are kept; there is no computation to port."""

    def __getitem__(self, launch):
        return self

    def __call__(self, Etheta, Ephi, DirectivityMap, az_range, el_range, wavelength):
        DirectivityMap[:len(az_range), :len(el_range)] = 0


GPUBeamformingMap = _ZeroMapKernel()


def MaximumDirectivityMap(Etheta, Ephi, source_coords, wavelength, az_range=None,
                          elev_range=None, az_index=None, elev_index=None, forming="Total",
                          total_solid_angle=(4 * np.pi), phase_resolution=24):
    """Best achieved directivity per command angle over equal-gain combining on
    Etheta, on Ephi and wavefront steering (beamforming.py:484-673).

    Returns (n_elev_index, n_az_index, 3) float64: Dtheta, Dphi, Dtot of the
    scheme that maximises the `forming` directivity.  As in the reference, the
    value kept for command (elev_inc, az_inc) is read at grid point
    (elev_inc, az_inc), the first scheme's directivity uses a 4 pi solid angle,
    and the power sum weights cells by the sine of their mean azimuth."""
    if az_range is None:
        az_range = np.linspace(-180.0, 180.0, 19)
    if elev_range is None:
        elev_range = np.linspace(-90.0, 90.0, 19)
    if elev_index is None:
        elev_index = np.linspace(0, len(elev_range) - 1, len(elev_range)).astype(int)
    if az_index is None:
        az_index = np.linspace(0, len(az_range) - 1, len(az_range)).astype(int)
    if forming not in ("Total", "Etheta", "Ephi"):
        raise ValueError("forming must be 'Total', 'Etheta' or 'Ephi'")
    Et = np.asarray(Etheta)
    Ep = np.asarray(Ephi)
    ne, n_el, n_az = Et.shape
    az_res, el_res = len(az_index), len(elev_index)
    pts = np.asarray(source_coords.points if hasattr(source_coords, "points") else source_coords)
    Et64 = Et.astype(np.complex64)
    Ep64 = Ep.astype(np.complex64)
    W, pstar = command_weights(Et64, Ep64, pts, wavelength, az_range, elev_range, az_index,
                               elev_index, phase_resolution)
    q = _power_weights(az_range, elev_range)
    power, ustar = _native.beam_power(W.reshape(-1, ne), Et64.reshape(ne, -1),
                                      Ep64.reshape(ne, -1), q, pstar.reshape(-1))
    power = power.reshape(az_res, el_res, 3)
    ustar = ustar.reshape(az_res, el_res, 3, 2).astype(np.float64)
    d_area = np.radians(az_range[1] - az_range[0]) * np.radians(elev_range[1] - elev_range[0])
    tsa = np.array([4 * np.pi, total_solid_angle, total_solid_angle])
    with np.errstate(divide="ignore", invalid="ignore"):
        Uav = power * d_area / tsa[None, None, :]
        Dt = ustar[..., 0] / Uav
        Dp = ustar[..., 1] / Uav
        Dtot = (ustar[..., 0] + ustar[..., 1]) / Uav
    comp = {"Total": Dtot, "Etheta": Dt, "Ephi": Dp}[forming]
    best = np.argmax(comp, axis=2)
    out = np.zeros((el_res, az_res, 3))
    ia, ie = np.meshgrid(np.arange(az_res), np.arange(el_res), indexing="ij")
    out[ie, ia, 0] = Dt[ia, ie, best]
    out[ie, ia, 1] = Dp[ia, ie, best]
    out[ie, ia, 2] = Dtot[ia, ie, best]
    return out
