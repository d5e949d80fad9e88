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
        normals = np.asarray(aperture_coords.point_data["Normals"]).astype(np.float32)
        return EM.calculate_conformalVectors(desired[0, :].reshape(1, 3), normals)
    with np.errstate(all="ignore"):
        row = np.real(desired[0, :]).astype(np.float32).reshape(1, 3)
    return np.repeat(row, num_sources, axis=0).astype(np.complex64)


def _point_table(points_list, normals_list, weights_list):
    pts = np.concatenate([np.asarray(p, np.float32).reshape(-1, 3) for p in points_list])
    nrm = np.concatenate([np.asarray(n, np.float32).reshape(-1, 3) for n in normals_list])
    w = np.concatenate([np.asarray(x, np.complex64).reshape(-1, 3) for x in weights_list])
    table = np.zeros(len(pts), dtype=scattering_t)
    table["px"], table["py"], table["pz"] = pts[:, 0], pts[:, 1], pts[:, 2]
    table["nx"], table["ny"], table["nz"] = nrm[:, 0], nrm[:, 1], nrm[:, 2]
    table["ex"], table["ey"], table["ez"] = w[:, 0], w[:, 1], w[:, 2]
    table["Electric"] = True
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

    if scattering != 1:
        raise NotImplementedError("calculate_scattering: scattering 0 and 1 are supported")
    scat = np.asarray(scatter_points.points).astype(np.float32)
    scat_n = np.asarray(scatter_points.point_data["Normals"]).astype(np.float32)
    P = scat.shape[0]
    w = _scatter_branch_weights(aperture_coords, desired_E_axis, project_vectors, S)
    if not elements:
        E = one_bounce_fields(src, w, snk, snk_n, scat, scat_n, tris9, alpha=alpha, beta=beta,
                              los=los)
        return E[:, 0], E[:, 1], E[:, 2]
    # per-source network: the reference's path list (S x P x K rows at most)
    src_n = np.asarray(aperture_coords.point_data["Normals"]).astype(np.float32)
    table = _point_table([src, snk, scat], [src_n, snk_n, scat_n],
                         [w, np.ones((K, 3)), np.ones((P, 3))])
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
    aperture = GF.compute_normals(GF.compute_areas(aperture))
    centroids = GF.cell_centroids(aperture)
    normals = np.asarray(aperture.cell_data["Normals"][-1])
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
