"""Time-domain model (lyceanem/models/time_domain.py) on the gfx950 kernels.

calculate_scattering keeps the reference signature and return values.  Paths
come from workchunkingv2 (GPU visibility); per-path amplitudes, arrival times,
the wake time and the placement of the excitation run in lem_time.hip."""
from __future__ import annotations

import numpy as np

from .. import _native
from ..base_types import triangles_as_array
from ..electromagnetics import empropagation as EM
from ..geometry import geometryfunctions as GF
from ..raycasting import rayfunctions as RF
from .frequency_domain import _point_table, _source_weights


def calculate_scattering(aperture_coords, sink_coords, excitation_function, antenna_solid,
                         desired_E_axis, scatter_points=None, wavelength=1.0, scattering=0,
                         elements=False, sampling_freq=1e9, los=False, num_samples=10000,
                         antenna_axes=np.eye(3), project_vectors=True, alpha=0.0,
                         beta=(np.pi * 2) / 1.0):
    """Time-domain response at the sinks (time_domain.py:14-525).

    Returns (Ex, Ey, Ez, WakeTimes).  elements=False: float64 (num_samples,)
    traces summed over sources and sinks, shifted to start at the wake time's
    sample (np.digitize on linspace(0, num_samples / sampling_freq,
    num_samples)).  elements=True: (num_sources, num_sinks, num_samples), where,
    as in the reference, row [e, k, :] holds the response summed over every
    sink to source e alone excited with its weight / num_sources (the same
    trace for every k), without the wake shift."""
    time_index = np.linspace(0, num_samples / sampling_freq, num_samples)
    src = np.asarray(aperture_coords.points).astype(np.float32)
    snk = np.asarray(sink_coords.points).astype(np.float32)
    S, K, N = src.shape[0], snk.shape[0], int(num_samples)
    tris9 = triangles_as_array(GF.mesh_conversion(antenna_solid))
    src_n = np.asarray(aperture_coords.point_data["Normals"]).astype(np.float32)
    snk_n = np.asarray(sink_coords.point_data["Normals"]).astype(np.float32)
    w = _source_weights(aperture_coords, np.asarray(desired_E_axis), project_vectors,
                        antenna_axes, S)
    if scattering == 0:
        scat = np.empty((0, 3), np.float32)
        scat_n = np.empty((0, 3), np.float32)
    else:
        scat = np.asarray(scatter_points.points).astype(np.float32)
        scat_n = np.asarray(scatter_points.point_data["Normals"]).astype(np.float32)
    P = scat.shape[0]
    full_index, _ = RF.workchunkingv2(src, snk, scat, tris9, scattering + 1, line_of_sight=los)
    exc = np.asarray(excitation_function, np.float64).reshape(-1)

    def table(weights):
        return _point_table([src, snk, scat], [src_n, snk_n, scat_n],
                            [weights, np.ones((K, 3)), np.ones((P, 3))])

    if not elements:
        Ex, Ey, Ez = (np.zeros(N, np.float64) for _ in range(3))
        if full_index.shape[0] == 0:
            return Ex, Ey, Ez, 1.0
        amp, _, wake, bins = EM.time_domain_paths(table(w), full_index, alpha, beta,
                                                  sampling_freq)
        trace = _native.time_trace(amp, bins, exc, N).cpu().numpy()
        wake_time = float(wake.item())
        wi = int(np.digitize(wake_time, time_index))
        for out, c in ((Ex, 0), (Ey, 1), (Ez, 2)):
            out[wi:] = trace[:N - wi, c] if wi < N else trace[:0, c]
        return Ex, Ey, Ez, wake_time

    Ex, Ey, Ez = (np.zeros((S, K, N), np.float64) for _ in range(3))
    if full_index.shape[0] == 0:
        return Ex, Ey, Ez, 1.0
    s_of_path, _ = EM.path_slots(full_index, S, K)
    if not np.any(np.imag(w)):
        # real weights: the excited element changes amplitudes only, never
        # arrival times, so one pass gives every element's trace (grouped by
        # source, amplitudes scaled by 1 / num_sources)
        amp, _, wake, bins = EM.time_domain_paths(table(w / S), full_index, alpha, beta,
                                                  sampling_freq)
        per = _native.time_map(amp, bins, s_of_path, S, exc, N).cpu().numpy()
        wake_time = float(wake.item())
        for out, c in ((Ex, 0), (Ey, 1), (Ez, 2)):
            out[:] = per[:, None, :, c]
        return Ex, Ey, Ez, wake_time
    wake_time = 1.0
    for e in range(S):
        we = np.zeros_like(w)
        we[e] = w[e] / S
        amp, _, wake, bins = EM.time_domain_paths(table(we), full_index, alpha, beta,
                                                  sampling_freq)
        tr = _native.time_trace(amp, bins, exc, N).cpu().numpy()
        wake_time = float(wake.item())
        Ex[e], Ey[e], Ez[e] = tr[None, :, 0], tr[None, :, 1], tr[None, :, 2]
    return Ex, Ey, Ez, wake_time
