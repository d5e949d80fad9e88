"""Visibility and chunking calls of the LyceanEM ray path, on the gfx950 kernels.

Same names, arguments and return values as the reference module
(lyceanem/raycasting/rayfunctions.py); the per-ray work runs in
`lem_raycast_pairs` / `lem_visibility_dense` (include/lem.h) instead of the
numba kernel kernel1Dv3.

Index conventions (unchanged): the unified point table is sources, then sinks,
then scattering points, and index rows are 1-based (0 = no point).
"""
from __future__ import annotations

import math

import numpy as np

from .. import _native
from .. import base_types
from ..mesh import triangle_cells
from ..utility import math_functions

EPSILON = 1e-7  # rayfunctions.py:19

# Bytes of device memory one ray costs in the chunked launchers (pair + flag).
_BYTES_PER_RAY = 2 * 4 + 1


def free_memory_bytes() -> int:
    """Free device memory used to size ray chunks.  Tests replace this function
    to force chunking; results must not depend on it."""
    import torch

    free, _ = torch.cuda.mem_get_info()
    return int(free)


def convertTriangles(triangle_object):
    """meshio-style triangle mesh -> triangle_t records (rayfunctions.py:531-557);
    float64 vertices are narrowed to float32 exactly as np.single does."""
    if triangle_object is None:
        return np.empty(0, dtype=base_types.triangle_t)
    tri = triangle_cells(triangle_object)
    v = np.asarray(triangle_object.points)
    out = np.empty(len(tri), dtype=base_types.triangle_t)
    flat = out.view(np.float32).reshape(-1, 9)
    for j in range(3):
        flat[:, 3 * j : 3 * j + 3] = v[tri[:, j]].astype(np.float32)
    return out


def create_model_index(source_index, sink_index, scattering_point_index):
    """Rows (source row..., target) for every source row and every target (sinks,
    then scattering points); rows whose last two columns match are dropped
    (rayfunctions.py:2624-2678)."""
    src = np.asarray(source_index)
    if src.ndim == 1:
        src = src.reshape(-1, 1)
    targets = np.asarray(sink_index).ravel()
    scat = np.asarray(scattering_point_index)
    if scat.shape[0] != 0:
        targets = np.append(targets, scat.ravel())
    n, m, c = src.shape[0], targets.shape[0], src.shape[1]
    io = np.zeros((n * m, c + 1), dtype=np.int32)
    if n * m:
        io[:, :c] = np.repeat(src, m, axis=0)
        io[:, c] = np.tile(targets, n)
    return io[io[:, -2] != io[:, -1], :]


def charge_rays_environment1Dv2(sources, sinks, environment_points, point_indexing):
    """Host ray payload (ray_t) for the index rows (rayfunctions.py:1790-1842).
    The GPU launchers charge rays on the device instead; this stays for callers
    that want the payload itself."""
    unified = np.concatenate([np.asarray(sources, np.float32), np.asarray(sinks, np.float32),
                              np.asarray(environment_points, np.float32).reshape(-1, 3)])
    idx = np.asarray(point_indexing)
    o = unified[idx[:, -2] - 1]
    d, L = math_functions.calc_dv_norm(o, unified[idx[:, -1] - 1])
    rays = np.empty(idx.shape[0], dtype=base_types.ray_t)
    rays["ox"], rays["oy"], rays["oz"] = o[:, 0], o[:, 1], o[:, 2]
    rays["dx"], rays["dy"], rays["dz"] = d[:, 0], d[:, 1], d[:, 2]
    rays["dist"] = L[:, 0]
    rays["intersect"] = False
    return rays


def rayHits1Dv2(ray_payload, point_indexing, sink_index):
    """Split unblocked rows into (filtered_index, final_index): rows ending on a
    sink are final, the rest continue (rayfunctions.py:1845-1878)."""
    inter = ray_payload["intersect"] if getattr(ray_payload, "dtype", None) is not None and \
        ray_payload.dtype.names else np.asarray(ray_payload, bool)
    idx = np.asarray(point_indexing)
    free = ~np.asarray(inter, bool)
    on_sink = np.isin(idx[:, -1], np.asarray(sink_index).ravel())
    return idx[free & ~on_sink, :], idx[free & on_sink, :]


def _tris9(environment_local):
    """(T, 9) float32 triangles of triangle_t records, a float array or meshes."""
    if hasattr(environment_local, "points") or (
            isinstance(environment_local, (list, tuple)) and environment_local
            and hasattr(environment_local[0], "points")):
        from ..geometry import geometryfunctions as GF

        environment_local = GF.mesh_conversion(environment_local)
    return base_types.triangles_as_array(environment_local)


def _cast_rows(unified, index, tris9):
    """Device raycast of index rows (last two columns = origin, target)."""
    idx = np.asarray(index)
    if idx.shape[0] == 0:
        return np.zeros(0, bool)
    pairs = np.ascontiguousarray(idx[:, -2:] - 1, dtype=np.int32)
    if pairs.min() < 0 or pairs.max() >= unified.shape[0]:
        raise IndexError("index rows point outside the unified point table")
    inter, _ = _native.raycast_pairs(unified, pairs, tris9, any_hit=True, want_dist=False)
    out = inter.cpu().numpy()
    if (out == 0xFF).any():
        raise IndexError("index rows point outside the unified point table")
    return out.astype(bool)


def _unified(sources, sinks, scattering_points):
    return np.ascontiguousarray(np.concatenate([
        np.asarray(sources, np.float32).reshape(-1, 3),
        np.asarray(sinks, np.float32).reshape(-1, 3),
        np.asarray(scattering_points, np.float32).reshape(-1, 3)]))


def launchRaycaster1Dv3(sources, sinks, scattering_points, io_indexing, environment_local):
    """Cast every row of io_indexing; returns (filtered_index, final_index,
    RAYS_CAST) (rayfunctions.py:1951-2017)."""
    unified = _unified(sources, sinks, scattering_points)
    S = np.asarray(sources).reshape(-1, 3).shape[0]
    K = np.asarray(sinks).reshape(-1, 3).shape[0]
    inter = _cast_rows(unified, io_indexing, _tris9(environment_local))
    sink_index = np.arange(S + 1, S + 1 + K).reshape(K, 1)
    filtered, final = rayHits1Dv2(inter, io_indexing, sink_index)
    return filtered, final, int(np.asarray(io_indexing).shape[0])


def chunkingRaycaster1Dv3(sources, sinks, scattering_points, filtered_index, environment_local,
                          terminate_flag):
    """Next wave from the surviving rows (rayfunctions.py:2102-2369): to the sinks
    only (terminate_flag) or to sinks and scattering points, split into chunks
    that fit half the free device memory."""
    S = np.asarray(sources).reshape(-1, 3).shape[0]
    K = np.asarray(sinks).reshape(-1, 3).shape[0]
    P = np.asarray(scattering_points).reshape(-1, 3).shape[0]
    sink_index = np.arange(S + 1, S + 1 + K).reshape(K, 1)
    scat_index = np.arange(S + K + 1, S + K + 1 + P).reshape(P, 1)
    unified = _unified(sources, sinks, scattering_points)
    tris9 = _tris9(environment_local)
    fi = np.asarray(filtered_index)
    ncol = fi.shape[1] + 1
    targets = K if terminate_flag else K + P
    ray_limit = max(1, int((0.5 * free_memory_bytes() - tris9.nbytes) // _BYTES_PER_RAY))
    nchunks = max(1, math.ceil(fi.shape[0] * max(targets, 1) / ray_limit))
    filtered2 = [np.empty((0, ncol), np.int32)]
    final2 = [np.empty((0, ncol), np.int32)]
    rays = 0
    for part in np.array_split(fi, nchunks):
        if terminate_flag:
            sub = create_model_index(part, sink_index, np.empty((0, 0), np.int32))
        else:
            sub = create_model_index(part, sink_index, scat_index)
        inter = _cast_rows(unified, sub, tris9)
        rays += sub.shape[0]
        f2, g2 = rayHits1Dv2(inter, sub, sink_index)
        filtered2.append(f2)
        final2.append(g2)
    return np.concatenate(filtered2), np.concatenate(final2), rays


def workchunkingv2(sources, sinks, scattering_points, environment, max_scatter,
                   line_of_sight=True):
    """Ray waves for max_scatter 1, 2 or 3 (rayfunctions.py:2681-2955); returns
    (full_index int32 (V, max_scatter+1), RAYS_CAST).

    max_scatter 2 keeps the reference's row layout: line-of-sight rows padded
    with a 0 column, then the (source, scatterer, sink) rows.  As in the
    reference (rayfunctions.py:2835-2861), when no scattering point is visible
    from any source the line-of-sight rows are not returned either.
    max_scatter 3 runs the reference's source chunks (np.linspace bounds,
    :2864-2867); per chunk the rows are [s, k, 0, 0], [s, p, k, 0] and
    [s, p1, p2, k], and RAYS_CAST counts the last chunk's rays as the
    reference's assignment does (:2936)."""
    S = np.asarray(sources).reshape(-1, 3).shape[0]
    K = np.asarray(sinks).reshape(-1, 3).shape[0]
    P = np.asarray(scattering_points).reshape(-1, 3).shape[0]
    tris9 = _tris9(environment)
    ray_limit = max(1, int((0.8 * free_memory_bytes() - tris9.nbytes) // _BYTES_PER_RAY))
    source_index = np.arange(1, S + 1).reshape(S, 1)
    sink_index = np.arange(S + 1, S + 1 + K).reshape(K, 1)
    scat_index = np.arange(S + K + 1, S + K + 1 + P).reshape(P, 1)
    if max_scatter == 1:
        io = create_model_index(source_index, sink_index, scat_index)
        parts = np.array_split(io, max(1, math.ceil(io.shape[0] / ray_limit))) \
            if io.shape[0] >= ray_limit else [io]
        full = [np.empty((0, 2), np.int32)]
        rays = 0
        for part in parts:
            _, fin, n = launchRaycaster1Dv3(sources, sinks, scattering_points, part, environment)
            full.append(fin)
            rays += n
        return np.concatenate(full).astype(np.int32), rays
    if max_scatter == 2:
        if line_of_sight:
            io = create_model_index(source_index, sink_index, scat_index)
        else:
            io = create_model_index(source_index, np.empty((0, 1)), scat_index)
        chunked = io.shape[0] >= ray_limit
        parts = np.array_split(io, max(1, math.ceil(io.shape[0] / ray_limit))) \
            if chunked else [io]
        full = [np.empty((0, 3), np.int32)]
        rays = 0
        for part in parts:
            fil, fin, n1 = launchRaycaster1Dv3(sources, sinks, scattering_points, part,
                                               environment)
            if fil.shape[0] == 0:
                if chunked:
                    # chunked branch (rayfunctions.py:2788-2831) keeps the LOS rows
                    full.append(np.concatenate([fin, np.zeros((fin.shape[0], 1), np.int32)],
                                               axis=1))
                rays += n1
                continue
            _, fin2, n2 = chunkingRaycaster1Dv3(sources, sinks, scattering_points, fil,
                                                environment, True)
            rays += n1 + n2
            full.append(np.concatenate(
                [np.concatenate([fin, np.zeros((fin.shape[0], 1), np.int32)], axis=1), fin2]))
        return np.concatenate(full).astype(np.int32), rays
    if max_scatter == 3:
        estimate = S * (K + P) + S * P * K * 3
        chunknum = max(2, min(S, math.ceil(estimate / 2e8)))
        bounds = np.linspace(0, S, chunknum, dtype=np.int32)
        full = [np.empty((0, 4), np.int32)]
        rays = 0
        for c in range(len(bounds) - 1):
            srows = source_index[bounds[c]:bounds[c + 1]]
            io = create_model_index(srows, sink_index if line_of_sight else np.empty((0, 1)),
                                    scat_index)
            fil, fin, n1 = launchRaycaster1Dv3(sources, sinks, scattering_points, io, environment)
            fil2, fin2, n2 = chunkingRaycaster1Dv3(sources, sinks, scattering_points, fil,
                                                   environment, False)
            _, fin3, n3 = chunkingRaycaster1Dv3(sources, sinks, scattering_points, fil2,
                                                environment, True)
            rays = n1 + n2 + n3
            fin = fin.reshape(-1, 2)
            t = np.concatenate([np.concatenate([fin, np.zeros((len(fin), 1), np.int32)], 1),
                                fin2.reshape(-1, 3)])
            full.append(np.concatenate([np.concatenate([t, np.zeros((len(t), 1), np.int32)], 1),
                                        fin3.reshape(-1, 4)]))
        return np.concatenate(full).astype(np.int32), rays
    raise NotImplementedError("workchunkingv2: max_scatter 1, 2 and 3 are supported (the "
                              "reference defines no other)")


def visiblespace(source_coords, source_normals, environment, vertex_area=0, az_range=None,
                 elev_range=None, shell_range=0.5):
    """Projected-area visibility pattern of an aperture (rayfunctions.py:316-427).

    Returns (visible_patterns (n_el, n_az) float32, sample shell mesh)."""
    from ..geometry import targets as TG

    if az_range is None:
        az_range = np.linspace(-180.0, 180.0, 19)
    if elev_range is None:
        elev_range = np.linspace(-90.0, 90.0, 19)
    pts = np.asarray(source_coords.points, np.float32)
    S = pts.shape[0]
    shell = TG.spherical_field(az_range, elev_range, outward_normals=True,
                               field_radius=shell_range)
    sinks = np.asarray(shell.points, np.float32)
    bits, _ = _native.visibility_dense(pts, sinks, _tris9(environment), any_hit=True)
    va = np.asarray(vertex_area, dtype=np.float64).ravel() if np.ndim(vertex_area) else \
        np.asarray([vertex_area], np.float64)
    va = np.where(np.isnan(va), 0.0, va)
    if va.size == 1:
        area = None if va[0] == 0 else np.full(S, va[0], np.float32)
    else:
        area = va.astype(np.float32)
    pattern = _native.projected_area(pts, np.asarray(source_normals, np.float32), area, sinks,
                                     bits).cpu().numpy()
    pattern = pattern.reshape(len(elev_range), len(az_range))
    shell.point_data["Projected_Area"] = pattern.reshape(-1, 1)
    return pattern, shell
