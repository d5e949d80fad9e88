"""Frequency-domain propagation of the LyceanEM path on the gfx950 kernels.

Same names and return values as the reference functions in
lyceanem/electromagnetics/empropagation.py; the per-path field runs in
`lem_field_paths` (include/lem.h) instead of the numba kernel freqdomainkernal.
"""
from __future__ import annotations

import numpy as np

from .. import _native


def targettingindex(full_index):
    """(source, sink) columns of every path and its bounce depth
    (empropagation.py:3394-3429): line-of-sight rows first, then one-bounce rows,
    each group in index order."""
    fi = np.asarray(full_index)
    scatter_index = np.ones(fi.shape[0], dtype=np.int64)
    if fi.shape[1] == 2:
        return fi, scatter_index
    if fi.shape[1] == 3:
        los = np.all(fi[:, 2:] == 0, axis=1)
        sc = np.all(fi[:, 2:] != 0, axis=1)
        depth = np.append(fi[los][:, [0, 1]], fi[sc][:, [0, 2]], axis=0)
        scatter_index[sc] = 2
        return depth, scatter_index
    if fi.shape[1] == 4:
        z2 = np.all(fi[:, 2:] == 0, axis=1)
        depth = fi[z2][:, [0, 1]]
        rem = fi[~z2]
        z3 = np.all(rem[:, 3:] == 0, axis=1)
        depth = np.append(depth, rem[z3][:, [0, 2]], axis=0)
        rem2 = rem[~z3]
        depth = np.append(depth, rem2[np.all(rem2[:, 3:] != 0, axis=1)][:, [0, 3]], axis=0)
        scatter_index[np.all(fi[:, 2:] != 0, axis=1)] = 3
        scatter_index[(fi[:, 2] != 0) & (fi[:, 3] == 0)] = 2
        return depth, scatter_index
    raise ValueError("full_index must have 2 to 4 columns")


def path_slots(full_index, source_num, sink_num):
    """0-based (source, sink) slot of every path row: the source column and the
    last non-zero column (the pairing targettingindex intends)."""
    fi = np.asarray(full_index, dtype=np.int64)
    last = fi[:, 1].copy()
    for c in range(2, fi.shape[1]):
        nz = fi[:, c] != 0
        last[nz] = fi[nz, c]
    s = fi[:, 0] - 1
    k = last - 1 - source_num
    return s, k


def EMGPUFreqDomain(source_num, sink_num, full_index, point_information, wavelength, alpha,
                    beta):
    """Scattering network (source_num, sink_num, 3) complex128 of every path in
    full_index (empropagation.py:2197-2410).

    point_information is the scattering_t table (base_types.py:89-119): points,
    normals and complex weights of sources, sinks and scattering points.  The
    paths of each (source, sink) slot are summed on the GPU in index order
    (Kahan-compensated float32), so the result does not depend on launch order;
    the reference adds them with float64 atomics in arbitrary order.  A path
    that starts at a magnetic point (Electric False) launches (w x d) / eta,
    eta = sqrt(Re permeability / Re permittivity) of that point
    (empropagation.py:1111-1146)."""
    pi = np.asarray(point_information)
    fi = np.asarray(full_index)
    S, K = int(source_num), int(sink_num)
    if fi.shape[0] == 0:
        return np.zeros((S, K, 3), dtype=np.complex128)
    pos, nrm, w, eta = _path_table(pi)
    s, k = path_slots(fi, S, K)
    if s.min() < 0 or s.max() >= S or k.min() < 0 or k.max() >= K:
        raise IndexError("full_index rows do not start on a source and end on a sink")
    slot = s * K + k
    order = np.argsort(slot, kind="stable")
    ss = slot[order]
    starts = np.flatnonzero(np.r_[True, ss[1:] != ss[:-1]])
    seg_start = np.append(starts, len(ss)).astype(np.int64)
    seg_slot = ss[starts].astype(np.int64)
    net = _native.field_paths(pos, nrm, w, np.ascontiguousarray(fi, np.int32), order, seg_start,
                              seg_slot, S, K, float(alpha), float(beta), eta=eta)
    return net.cpu().numpy()


def vector_mapping(local_E_vector, point_normal, rotation_matrix):
    """Local (u, v, n) excitation -> global vector (empropagation.py:3692-3762):
    n is the rotated point normal, u = z x n / |z x n| (x x n when n is along z),
    v = n x u / |u x n|, and the result is local_E_vector @ [u; v; n]."""
    dt = np.asarray(local_E_vector).dtype
    n = np.matmul(np.asarray(point_normal).astype(dt), np.asarray(rotation_matrix))
    axes = np.eye(3)
    uvn = np.zeros((3, 3), dtype=dt)
    uvn[2, :] = n
    z_orth = np.linalg.norm(np.cross(axes[:, 2], n))
    if abs(z_orth) == 0:
        c = np.cross(axes[0, :], n)
    else:
        c = np.cross(axes[2, :], n)
    uvn[0, :] = c / np.linalg.norm(c)
    uvn[1, :] = np.cross(n, uvn[0, :]) / np.linalg.norm(np.cross(uvn[0, :], n))
    return np.matmul(local_E_vector, uvn)


def calculate_conformalVectors(desired_E_vector, source_normals, antenna_axes):
    """Per-point conformal excitation vectors (empropagation.py:3807-3849);
    antenna_axes is required, as in the reference."""
    normals = np.asarray(source_normals)
    out = np.zeros((normals.shape[0], 3), dtype=np.complex64)
    dv = np.asarray(desired_E_vector)
    ax = np.asarray(antenna_axes).astype(np.complex64)
    per_point = dv.ndim == 2 and dv.shape[0] == normals.shape[0]
    for i in range(normals.shape[0]):
        e = dv[i, :] if per_point else dv.ravel()
        out[i, :] = vector_mapping(e, normals[i, :], ax)
    return out


def _path_table(point_information):
    """Points, normals, weights and per-point impedances (None when every point
    is electric; 0 marks an electric point) of a scattering_t table."""
    pi = np.asarray(point_information)
    pos = np.stack([pi["px"], pi["py"], pi["pz"]], axis=1).astype(np.float32)
    nrm = np.stack([pi["nx"], pi["ny"], pi["nz"]], axis=1).astype(np.float32)
    w = np.stack([pi["ex"], pi["ey"], pi["ez"]], axis=1).astype(np.complex64)
    eta = None
    magnetic = ~pi["Electric"].astype(bool) if pi.shape[0] else np.zeros(0, bool)
    if magnetic.any():
        mu = pi["permeability"].real.astype(np.float64)
        eps = pi["permittivity"].real.astype(np.float64)
        with np.errstate(divide="ignore", invalid="ignore"):
            z = np.sqrt(mu / eps)
        bad = magnetic & ~(np.isfinite(z) & (z > 0))
        if bad.any():
            raise ValueError("magnetic points need a positive real impedance "
                             "sqrt(Re permeability / Re permittivity)")
        eta = np.where(magnetic, z, 0.0).astype(np.float32)
    return pos, nrm, w, eta


def time_domain_paths(point_information, full_index, alpha, beta, sampling_freq):
    """(amp (V, 3), arrival (V,), wake (1,), bins (V,)) device tensors of every
    row of full_index: timedomainkernal's per-path part and start samples; a
    magnetic first point launches Re(w) x d / eta (empropagation.py:1395-1419)."""
    pos, nrm, w, eta = _path_table(point_information)
    fi = np.ascontiguousarray(full_index, np.int32)
    amp, arrival = _native.time_paths(pos, nrm, w, fi, alpha, beta, eta=eta)
    wake = _native.time_wake(arrival)
    bins = _native.time_bins(arrival, wake, sampling_freq)
    return amp, arrival, wake, bins


def TimeDomainv3(source_num, sink_num, point_informationv2, full_index, scattering_coefficient,
                 wavelength, excitation_signal, sampling_freq, num_samples, alpha, beta):
    """Time-domain response of every (source, sink) pair (empropagation.py:2910-3170).

    Returns (time_map, wake_time): time_map float64 (source_num, sink_num,
    num_samples, 3) holds, per pair, the excitation placed at each path's start
    sample (arrival - wake) // (1 / sampling_freq) and scaled by the path's
    polarised amplitude; wake_time is the earliest arrival (at most 1 s).  The
    paths of a pair are added in index order (the reference uses float64
    atomics), and the wake time is the minimum over every path (the
    reference's 1 GB source chunking, with a wake time per chunk re-aligned by
    rounding, is not reproduced).  scattering_coefficient and wavelength are
    unused, as in the reference."""
    S, K, N = int(source_num), int(sink_num), int(num_samples)
    fi = np.asarray(full_index)
    if fi.shape[0] == 0:
        return np.zeros((S, K, N, 3), np.float64), 1.0
    amp, _, wake, bins = time_domain_paths(point_informationv2, fi, alpha, beta, sampling_freq)
    s, k = path_slots(fi, S, K)
    if s.min() < 0 or s.max() >= S or k.min() < 0 or k.max() >= K:
        raise IndexError("full_index rows do not start on a source and end on a sink")
    tm = _native.time_map(amp, bins, s * K + k, S * K, excitation_signal, N)
    return tm.cpu().numpy().reshape(S, K, N, 3), float(wake.item())
