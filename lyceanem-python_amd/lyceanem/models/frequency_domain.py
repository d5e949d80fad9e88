"""Frequency-domain model API of LyceanEM on the gfx950 kernels.

`calculate_farfield`, `calculate_scattering` and `aperture_projection` keep the
reference signatures and return values (lyceanem/models/frequency_domain.py).

Line of sight (scattering=0) runs as two fused device passes: a dense
source x sink visibility bitmask (`lem_visibility_dense`) and the masked field
kernel (`lem_field_los`), so the S*K ray index and payload the reference builds
on the host never exist.  One bounce (scattering=1) summed over sources is
evaluated leg by leg (source -> scatterers, scatterers -> sinks, plus line of
sight), which equals the reference's per-path sum by linearity; the per-source
network (elements=True) follows the reference's path list (workchunkingv2 ->
EMGPUFreqDomain), both on the device kernels.
"""
from __future__ import annotations

import copy

import numpy as np

from .. import _native
from ..electromagnetics import empropagation as EM
from ..geometry import geometryfunctions as GF
from ..raycasting import rayfunctions as RF
from ..base_types import scattering_t, triangles_as_array

EPSILON = 1e-7


def _source_weights(aperture_coords, desired_E_axis, project_vectors, antenna_axes,
                    num_sources):
    """Source excitation of the line-of-sight branch (frequency_domain.py:339-380):
    the desired vector per source (projected if asked, repeated if given once),
    times each source's Area, stored as complex64."""
    normals = np.asarray(aperture_coords.point_data["Normals"])
    desired = np.asarray(desired_E_axis)
    if project_vectors:
        vec = EM.calculate_conformalVectors(desired, normals, antenna_axes)
    elif desired.shape[0] == normals.shape[0]:
        vec = copy.deepcopy(desired)
    else:
        vec = np.repeat(desired.reshape(1, 3).astype(np.complex64), num_sources, axis=0)
    vec = vec * np.asarray(aperture_coords.point_data["Area"]).reshape(-1, 1)
    return np.asarray(vec).astype(np.complex64)


def _scatter_branch_weights(aperture_coords, desired_E_axis, project_vectors, num_sources):
    """Source excitation of the scattering > 0 branch (frequency_domain.py:452-485):
    the reference keeps only row 0 of desired_E_axis, cast to float32 (the
    imaginary part is dropped), and does not multiply by Area."""
    desired = np.atleast_2d(np.asarray(desired_E_axis))
    if project_vectors:
        # frequency_domain.py:456-459 passes two of calculate_conformalVectors'
        # three arguments (empropagation.py:3807), which raises there
        raise TypeError("calculate_scattering(scattering > 0, project_vectors=True): the "
                        "reference calls calculate_conformalVectors without antenna_axes "
                        "(frequency_domain.py:456-459) and raises; project the vectors with "
                        "empropagation.calculate_conformalVectors and pass them as "
                        "desired_E_axis instead")
    with np.errstate(all="ignore"):
        row = np.real(desired[0, :]).astype(np.float32).reshape(1, 3)
    return np.repeat(row, num_sources, axis=0).astype(np.complex64)


def _point_table(points_list, normals_list, weights_list, permittivity=8.8541878176e-12,
                 permeability=1.25663706212e-6):
    """scattering_t table of electric points with the given free-space
    constants, as the reference fills point_informationv2
    (frequency_domain.py:401-405)."""
    pts = np.concatenate([np.asarray(p, np.float32).reshape(-1, 3) for p in points_list])
    nrm = np.concatenate([np.asarray(n, np.float32).reshape(-1, 3) for n in normals_list])
    w = np.concatenate([np.asarray(x, np.complex64).reshape(-1, 3) for x in weights_list])
    table = np.zeros(len(pts), dtype=scattering_t)
    table["px"], table["py"], table["pz"] = pts[:, 0], pts[:, 1], pts[:, 2]
    table["nx"], table["ny"], table["nz"] = nrm[:, 0], nrm[:, 1], nrm[:, 2]
    table["ex"], table["ey"], table["ez"] = w[:, 0], w[:, 1], w[:, 2]
    table["Electric"] = True
    table["permittivity"] = permittivity
    table["permeability"] = permeability
    return table


def calculate_scattering(aperture_coords, sink_coords, antenna_solid, desired_E_axis,
                         scatter_points=None, wavelength=1.0, scattering=0, elements=False,
                         los=True, project_vectors=False, antenna_axes=np.eye(3), multiE=False,
                         alpha=0.0, beta=(np.pi * 2) / 1.0, cuda=False,
                         acceleration_structure=None, chunks=1,
                         permittivity=8.8541878176e-12, permeability=1.25663706212e-6):
    """Fields at the sinks from the sources, with blocking by antenna_solid
    (frequency_domain.py:267-926).

    Returns (Ex, Ey, Ez): complex128 (K,) summed over sources, or (S, K) per
    source with elements=True.  cuda=True follows the reference's tile
    (acceleration-structure) branch."""
    if multiE:
        raise NotImplementedError(
            "multiE: the reference's multiE branch calls calculate_conformalVectors with two "
            "of its three arguments (frequency_domain.py:612-617) and cannot run")
    if cuda:
        from .acceleration_structures import calculate_scattering_tiles_branch

        return calculate_scattering_tiles_branch(
            aperture_coords, sink_coords, antenna_solid, desired_E_axis, scatter_points,
            wavelength, scattering, elements, project_vectors, antenna_axes, alpha, beta,
            acceleration_structure, chunks)
    src = np.asarray(aperture_coords.points).astype(np.float32)
    snk = np.asarray(sink_coords.points).astype(np.float32)
    S, K = src.shape[0], snk.shape[0]
    tris9 = triangles_as_array(GF.mesh_conversion(antenna_solid))
    snk_n = np.asarray(sink_coords.point_data["Normals"]).astype(np.float32)

    if scattering == 0:
        w = _source_weights(aperture_coords, desired_E_axis, project_vectors, antenna_axes, S)
        bits, _ = _native.visibility_dense(src, snk, tris9, any_hit=True)
        out = _native.field_los(src, w, snk, snk_n, hit_bits=bits, alpha=alpha, beta=beta,
                                elements=elements)
        out = out.cpu().numpy().astype(np.complex128)
        return out[..., 0], out[..., 1], out[..., 2]

    if scattering not in (1, 2):
        # workchunkingv2 defines max_scatter 1 to 3 only (rayfunctions.py:2745-2953)
        raise NotImplementedError("calculate_scattering: scattering 0, 1 and 2 are supported")
    scat = np.asarray(scatter_points.points).astype(np.float32)
    scat_n = np.asarray(scatter_points.point_data["Normals"]).astype(np.float32)
    P = scat.shape[0]
    w = _scatter_branch_weights(aperture_coords, desired_E_axis, project_vectors, S)
    if not elements:
        if scattering == 1:
            E = one_bounce_fields(src, w, snk, snk_n, scat, scat_n, tris9, alpha=alpha,
                                  beta=beta, los=los)
        else:
            E = two_bounce_fields(src, w, snk, snk_n, scat, scat_n, tris9, alpha=alpha,
                                  beta=beta, los=los)
        return E[:, 0], E[:, 1], E[:, 2]
    # per-source network: the reference's path list (S x P x K rows at most)
    src_n = np.asarray(aperture_coords.point_data["Normals"]).astype(np.float32)
    table = _point_table([src, snk, scat], [src_n, snk_n, scat_n],
                         [w, np.ones((K, 3)), np.ones((P, 3))], permittivity, permeability)
    full_index, _ = RF.workchunkingv2(src, snk, scat, tris9, scattering + 1, line_of_sight=los)
    net = EM.EMGPUFreqDomain(S, K, full_index, table, wavelength, alpha, beta)
    return net[:, :, 0], net[:, :, 1], net[:, :, 2]


# Above this many triangles visibility is cast through a uniform grid (DDA)
# instead of testing every triangle; both give the same masks.
GRID_MIN_TRIANGLES = 4096


def _visibility_bits(src, snk, tris9, grid):
    if grid is not None:
        bits, _ = grid.visibility(src, snk, any_hit=True)
    else:
        bits, _ = _native.visibility_dense(src, snk, tris9, any_hit=True)
    return bits


def one_bounce_fields(src, src_w, snk, snk_n, scat, scat_n, tris9, scat_w=None, alpha=0.0,
                      beta=2 * np.pi, los=True, grid=None):
    """Source-summed fields of every line-of-sight and one-bounce path, as
    calculate_scattering(scattering=1) gets from workchunkingv2 (max_scatter 2)
    and EMGPUFreqDomain, evaluated leg by leg:

        H_p = P(n_p) sum_s vis(s,p) L_sp P(d_sp) w_s * w_p      (leg 1, at scatterers)
        E_k = P(n_k) sum_s vis(s,k) L_sk P(d_sk) w_s            (line of sight)
            + sum_p vis(p,k) L_pk P(d_pk) H_p                   (leg 2)

    P(n) = I - n n^T; the line-of-sight rows of a 3-column path index are
    projected onto the sink tangent plane (empropagation.py:1150-1161).
    scat_w: per-scatterer weight (the reference fixes it to 1,
    frequency_domain.py:513-515).  As in workchunkingv2's unchunked branch
    (rayfunctions.py:2835-2861), no visible scatterer means no paths at all.
    Returns (K, 3) complex128."""
    tris9 = np.asarray(tris9, np.float32).reshape(-1, 9)
    T = tris9.shape[0]
    if grid is None and T >= GRID_MIN_TRIANGLES:
        grid = _native.Grid(tris9)
    P = len(scat)
    K = len(snk)
    if scat_w is None:
        scat_w = np.ones((P, 3), np.complex64)
    if P == 0:
        return np.zeros((K, 3), np.complex128)
    S = len(src)
    if grid is not None:
        # sources are the (compact) aperture, scatterers are spread over the
        # scene: cast 64 neighbouring sources to one scatterer per wave
        leg1 = {}
        H = _native.field_los_grid_sink_major(grid, src, src_w, scat, scat_n, snk_w=scat_w,
                                              alpha=alpha, beta=beta, sink_tangent=True,
                                              stats=leg1)
        blocked = leg1["blocked"]
    else:
        bits = _visibility_bits(src, scat, tris9, None)
        blocked = _native.count_blocked(bits)
        H = _native.field_los(src, src_w, scat, scat_n, hit_bits=bits, snk_w=scat_w,
                              alpha=alpha, beta=beta, sink_tangent=True)
    if blocked == S * P:
        # every source -> scatterer ray is blocked: workchunkingv2 returns no
        # rows, line of sight included (rayfunctions.py:2835-2861)
        return np.zeros((K, 3), np.complex128)
    if grid is not None:
        # leg 2: one scatterer to 64 neighbouring sinks per wave; line of sight:
        # 64 neighbouring aperture points to one sink per wave
        E = _native.field_los_grid_masked(grid, scat, H, snk, snk_n, alpha=alpha, beta=beta,
                                          sink_major=False)
        if los:
            E = E + _native.field_los_grid_masked(grid, src, src_w, snk, snk_n, alpha=alpha,
                                                  beta=beta, sink_tangent=True, sink_major=True)
    else:
        bits = _visibility_bits(scat, snk, tris9, None)
        E = _native.field_los(scat, H, snk, snk_n, hit_bits=bits, alpha=alpha, beta=beta)
        if los:
            bits = _visibility_bits(src, snk, tris9, None)
            E = E + _native.field_los(src, src_w, snk, snk_n, hit_bits=bits, alpha=alpha,
                                      beta=beta, sink_tangent=True)
    return E.cpu().numpy().astype(np.complex128)


def _project_plane(E, n):
    """E - n (n . E) for a unit-normalised n: the tangent-plane projector
    freqdomainkernal applies at a point (sourcelaunchtransformGPU with the
    point's normal, empropagation.py:178-249, 1147-1156)."""
    n = np.asarray(n, np.float64).reshape(3)
    n = n / np.linalg.norm(n)
    nt = _native.to_dev(n).to(E.dtype)
    return E - nt[None, :] * (E * nt[None, :]).sum(-1, keepdim=True)


def two_bounce_fields(src, src_w, snk, snk_n, scat, scat_n, tris9, scat_w=None, alpha=0.0,
                      beta=2 * np.pi, los=True, grid=None):
    """Source-summed fields of every path calculate_scattering(scattering=2)
    builds -- workchunkingv2 with max_scatter 3 (rayfunctions.py:2863-2953):
    rows [s, k, 0, 0], [s, p, k, 0] and [s, p1, p2, k] -- through
    freqdomainkernal (empropagation.py:1083-1265), evaluated leg by leg:

        H_p = P(n_p) (sum_s vis(s,p) L_sp P(d_sp) w_s * w_p)            at scatterers
        G_q = P(n_q) (sum_{p != q} vis(p,q) L_pq P(d_pq) H_p * w_q)     second bounce
        E_k = P(n_N) P(n_k) sum_s vis(s,k) L_sk P(d_sk) w_s             line of sight
            + P(n_k) sum_p vis(p,k) L_pk P(d_pk) H_p                    one bounce
            + sum_q vis(q,k) L_qk P(d_qk) G_q                           two bounces

    P(n) = I - n n^T.  In a 4-column row the kernel visits columns 1 and 2 as
    intermediate points: the sink of a one-bounce row sits in column 2 and is
    projected onto its tangent plane, and the zero column 2 of a line-of-sight
    row reads point -1 (numba wraps the negative index), so those rows are also
    projected with n_N, the normal of the last point of the table (the last
    scatterer).  Unlike max_scatter 2, the line of sight is kept when no
    scatterer is visible.  Returns (K, 3) complex128."""
    tris9 = np.asarray(tris9, np.float32).reshape(-1, 9)
    T = tris9.shape[0]
    if grid is None and T >= GRID_MIN_TRIANGLES:
        grid = _native.Grid(tris9)
    P, K = len(scat), len(snk)
    if K == 0:
        return np.zeros((0, 3), np.complex128)
    if scat_w is None:
        scat_w = np.ones((P, 3), np.complex64)
    # the last row of the point table (sources, sinks, scatterers)
    n_last = scat_n[-1] if P else snk_n[-1]

    def leg(a, a_w, b, b_n, b_w=None, tangent=False, sink_major=True, from_sink=False,
            same=False):
        if grid is not None:
            return _native.field_los_grid_masked(grid, a, a_w, b, b_n, snk_w=b_w, alpha=alpha,
                                                 beta=beta, sink_tangent=tangent,
                                                 sink_major=sink_major, from_sink=from_sink,
                                                 same_points=same)
        bits = _visibility_bits(a, b, tris9, None)
        if same:
            _native.block_same_point(bits, 0, len(a))
        return _native.field_los(a, a_w, b, b_n, hit_bits=bits, snk_w=b_w, alpha=alpha,
                                 beta=beta, sink_tangent=tangent)

    E = None
    if los:
        E = _project_plane(leg(src, src_w, snk, snk_n, tangent=True), n_last)
    if P:
        # array -> scatterers: 64 neighbouring elements to one scatterer per wave
        H = leg(src, src_w, scat, scat_n, b_w=scat_w, tangent=True, from_sink=True)
        # scatterers -> scatterers: one scatterer to 64 neighbouring ones per wave
        G = leg(scat, H, scat, scat_n, b_w=scat_w, tangent=True, sink_major=False, same=True)
        E1 = leg(scat, H, snk, snk_n, tangent=True, sink_major=False)
        E2 = leg(scat, G, snk, snk_n, sink_major=False)
        E = E1 + E2 if E is None else E + E1 + E2
    if E is None:
        return np.zeros((K, 3), np.complex128)
    return E.cpu().numpy().astype(np.complex128)


def calculate_farfield(aperture_coords, antenna_solid, desired_E_axis, az_range=None,
                       el_range=None, scatter_points=None, wavelength=1.0, farfield_distance=2.0,
                       scattering=0, elements=False, los=True, project_vectors=False, cuda=False,
                       antenna_axes=np.eye(3), alpha=0.0, beta=(np.pi * 2) / 1.0):
    """E-theta / E-phi on an (el, az) grid (frequency_domain.py:94-192).

    As in the reference, farfield_distance reaches spherical_field's
    `outward_normals` slot (frequency_domain.py:159 vs targets.py:448), so the
    sinks sit on the unit sphere with outward normals whatever its value."""
    from ..geometry.targets import spherical_field

    if az_range is None:
        az_range = np.linspace(-180.0, 180.0, 19)
    if el_range is None:
        el_range = np.linspace(-90.0, 90.0, 19)
    sink_coords = spherical_field(az_range, el_range, farfield_distance)
    Ex, Ey, Ez = calculate_scattering(
        aperture_coords, sink_coords, antenna_solid, desired_E_axis,
        scatter_points=scatter_points, wavelength=wavelength, scattering=scattering,
        elements=elements, los=los, project_vectors=project_vectors, antenna_axes=antenna_axes,
        multiE=False, cuda=cuda, alpha=alpha, beta=beta)
    th = sink_coords.point_data["theta_(Radians)"]
    ph = sink_coords.point_data["phi_(Radians)"]
    etheta = Ex * np.cos(ph) * np.cos(th) + Ey * np.sin(ph) * np.cos(th) - Ez * np.sin(th)
    ephi = -Ex * np.sin(ph) + Ey * np.cos(ph)
    return etheta, ephi


def aperture_projection(aperture, environment=None, wavelength=1.0, az_range=None,
                        elev_range=None, farfield_distance=2.0):
    """Maximum directivity envelope of an aperture from its visible projected area
    (frequency_domain.py:15-91)."""
    if az_range is None:
        az_range = np.linspace(-180.0, 180.0, 19)
    if elev_range is None:
        elev_range = np.linspace(-90.0, 90.0, 19)
    blocking = GF.mesh_conversion(aperture if environment is None else environment)
    # the reference's order (frequency_domain.py:63-80): the centroid cloud takes
    # the aperture's own cell data (its "Area" weights the projected area), the
    # normals come from the area / normal helpers' copy of the mesh
    centroids = GF.cell_centroids(aperture)
    if "Area" not in centroids.point_data:
        raise KeyError("aperture_projection weights each triangle by the aperture's own cell "
                       "data 'Area' (frequency_domain.py:63-80); compute it first, e.g. with "
                       "geometryfunctions.compute_areas")
    meshed = GF.compute_normals(GF.compute_areas(aperture))
    normals = np.asarray(meshed.cell_data["Normals"][GF.locate_cell_index(meshed)])
    centroids.points = centroids.points + EPSILON * normals
    pattern, pcd = RF.visiblespace(centroids, normals, blocking,
                                   vertex_area=centroids.point_data["Area"], az_range=az_range,
                                   elev_range=elev_range, shell_range=farfield_distance)
    envelope = (4 * np.pi * pattern) / (wavelength ** 2)
    pcd.point_data["Directivity_Envelope"] = (4 * np.pi * pcd.point_data["Projected_Area"]) / (
        wavelength ** 2)
    with np.errstate(divide="ignore"):
        pcd.point_data["Directivity_Envelope_(dBi)"] = 10 * np.log10(
            pcd.point_data["Directivity_Envelope"])
    return envelope.astype(np.float32), pcd
