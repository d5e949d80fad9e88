"""Acceleration-structure classes of the native branch
(lyceanem/models/acceleration_structures.py), on the gfx950 kernels.

`Tile_acceleration_structure` keeps the reference constructor, attributes and
`calculate_scattering` signature; its tiles are binned by `lyceanem.em` and the
visibility runs through the uniform grid with the native predicate.
`Brute_Force_acceleration_structure` works (the reference's wrapper uses
attributes it never sets and float slice bounds, acceleration_structures.py:
146-172)."""
from __future__ import annotations

import numpy as np

from .. import em
from ..mesh import triangle_cells


def source_chunks(S, chunk_count):
    """Source row ranges of the chunked tile calls (acceleration_structures.py:
    76-113): chunks of floor(S / chunk_count) rows, the last one running to S.
    With self_to_self=False the native kernel skips pairs whose chunk-local source
    row equals the sink row (raycasting_accelerated.cuh:61-66), so these bounds
    decide which pairs are skipped."""
    step = S // chunk_count
    out, a, b = [], 0, step
    for i in range(chunk_count):
        if i == chunk_count - 1:
            b = S
        out.append((a, b))
        a, b = b, b + step
    return out


class Tile_acceleration_structure:
    """(y, z) tile binning of a blocking mesh (acceleration_structures.py:8-115)."""

    def __init__(self, blocking_mesh, n_cells):
        pts = np.asarray(blocking_mesh.points)
        self.triangle_verticies = np.ascontiguousarray(pts)
        self.max_x = np.max(pts[:, 0]) + abs(np.max(pts[:, 0]) * 0.001)
        self.min_x = np.min(pts[:, 0]) - abs(np.min(pts[:, 0]) * 0.001)
        self.max_y = np.max(pts[:, 1]) + abs(np.max(pts[:, 1]) * 0.001)
        self.min_y = np.min(pts[:, 1]) - abs(np.min(pts[:, 1]) * 0.001)
        self.max_z = np.max(pts[:, 2]) + abs(np.max(pts[:, 2]) * 0.001)
        self.min_z = np.min(pts[:, 2]) - abs(np.min(pts[:, 2]) * 0.001)
        diff_y = self.max_y - self.min_y
        diff_z = self.max_z - self.min_z
        self.tile_size = min(diff_y, diff_z) / n_cells
        self.y_cells_count = int(np.ceil(diff_y / self.tile_size))
        self.z_cells_count = int(np.ceil(diff_z / self.tile_size))
        tri = triangle_cells(blocking_mesh)
        self.bin_counts = em.bin_counts_to_numpy(pts, tri, self.y_cells_count,
                                                 self.z_cells_count, self.min_y, self.tile_size,
                                                 self.min_z)
        self.binned_triangles_count = int(np.sum(self.bin_counts))
        self.binned_triangles = em.bin_triangles_to_numpy(
            pts, tri, self.y_cells_count, self.z_cells_count, self.min_y, self.tile_size,
            self.min_z, self.bin_counts, self.binned_triangles_count)

    def calculate_scattering(self, source_mesh, sink_mesh, alpha, beta, wavelength, self_to_self,
                             chunk_count=1):
        """(S, K, 3) complex64 per-pair network (acceleration_structures.py:72-115)."""
        assert chunk_count > 0, "chunk_count must be greater than 0"
        S = source_mesh.points.shape[0]
        K = sink_mesh.points.shape[0]
        out = np.zeros((S, K, 3), dtype=np.complex64)
        for a, b in source_chunks(S, chunk_count):
            if b <= a:
                continue
            ex = np.asarray(source_mesh.point_data["ex"])[a:b]
            ey = np.asarray(source_mesh.point_data["ey"])[a:b]
            ez = np.asarray(source_mesh.point_data["ez"])[a:b]
            arr = em.calculate_scattering_tiles(
                np.asarray(source_mesh.points)[a:b], np.asarray(sink_mesh.points),
                self.triangle_verticies, wavelength, ex.real, ex.imag, ey.real, ey.imag, ez.real,
                ez.imag,
                np.append(np.asarray(source_mesh.point_data["Normals"])[a:b],
                          np.asarray(sink_mesh.point_data["Normals"]), axis=0),
                self.min_x, self.max_x, self.min_y, self.max_y, self.min_z, self.max_z,
                self.tile_size, self.bin_counts, self.binned_triangles, self.y_cells_count,
                self.z_cells_count, self.binned_triangles_count, alpha, beta, self_to_self)
            out[a:b] = np.ascontiguousarray(arr).view(np.complex64).reshape(b - a, K, 3)
        return out


class Brute_Force_acceleration_structure:
    """Every triangle tested (acceleration_structures.py:118-184)."""

    def __init__(self, blocking_mesh):
        self.triangle_verticies = np.ascontiguousarray(blocking_mesh.points)
        self.triangles = triangle_cells(blocking_mesh)

    def calculate_scattering(self, source_mesh, sink_mesh, alpha, beta, wavelength, self_to_self,
                             chunk_count=1):
        assert chunk_count > 0, "chunk_count must be greater than 0"
        S = source_mesh.points.shape[0]
        K = sink_mesh.points.shape[0]
        ex = np.asarray(source_mesh.point_data["ex"])
        ey = np.asarray(source_mesh.point_data["ey"])
        ez = np.asarray(source_mesh.point_data["ez"])
        arr = em.calculate_scattering_brute_force(
            np.asarray(source_mesh.points), np.asarray(sink_mesh.points), self.triangles,
            self.triangle_verticies, wavelength, ex.real, ex.imag, ey.real, ey.imag, ez.real,
            ez.imag,
            np.append(np.asarray(source_mesh.point_data["Normals"]),
                      np.asarray(sink_mesh.point_data["Normals"]), axis=0),
            alpha, beta, self_to_self)
        return np.ascontiguousarray(arr).view(np.complex64).reshape(S, K, 3)


def calculate_scattering_tiles_branch(aperture_coords, sink_coords, antenna_solid,
                                      desired_E_axis, scatter_points, wavelength, scattering,
                                      elements, project_vectors, antenna_axes, alpha, beta,
                                      acceleration_structure, chunks):
    """calculate_scattering(cuda=True) (frequency_domain.py:719-926) with
    calculate_scattering_cuda (frequency_domain.py:195-264)."""
    from ..electromagnetics import empropagation as EM
    from ..geometry import geometryfunctions as GF
    from ..mesh import Mesh

    S = len(np.asarray(aperture_coords.points))
    if acceleration_structure is None:
        meshes = GF._as_meshes(antenna_solid)
        pts, tris, off = [], [], 0
        for m in meshes:
            p = np.asarray(m.points)
            pts.append(p)
            tris.append(triangle_cells(m) + off)
            off += len(p)
        env = Mesh(points=np.concatenate(pts), cells=[("triangle", np.concatenate(tris))])
        tile = Tile_acceleration_structure(env, 1)
    else:
        tile = acceleration_structure
    normals = np.asarray(aperture_coords.point_data["Normals"])
    desired = np.asarray(desired_E_axis)
    if project_vectors:
        vec = EM.calculate_conformalVectors(desired, normals, antenna_axes)
    elif desired.shape[0] == normals.shape[0]:
        vec = desired.copy()
    else:
        vec = np.repeat(desired.reshape(1, 3).astype(np.complex64), S, axis=0)
    vec = vec * np.asarray(aperture_coords.point_data["Area"]).reshape(-1, 1)
    aperture_coords.point_data["ex"] = np.ascontiguousarray(vec[:, 0])
    aperture_coords.point_data["ey"] = np.ascontiguousarray(vec[:, 1])
    aperture_coords.point_data["ez"] = np.ascontiguousarray(vec[:, 2])
    net = tile.calculate_scattering(aperture_coords, sink_coords, alpha, beta, wavelength,
                                    self_to_self=False, chunk_count=chunks)
    if scattering >= 1:
        src_scat = tile.calculate_scattering(aperture_coords, scatter_points, alpha, beta,
                                             wavelength, self_to_self=False, chunk_count=chunks)
        ones = np.ones(S)
        scatter_points.point_data["ex"] = np.dot(ones, src_scat[:, :, 0])
        scatter_points.point_data["ey"] = np.dot(ones, src_scat[:, :, 1])
        scatter_points.point_data["ez"] = np.dot(ones, src_scat[:, :, 2])
        scat_sink = tile.calculate_scattering(scatter_points, sink_coords, alpha, beta,
                                              wavelength, self_to_self=False,
                                              chunk_count=chunks)
        if scattering == 1:
            for c in range(3):
                net[:, :, c] += np.dot(src_scat[:, :, c], scat_sink[:, :, c])
    if elements:
        return net[:, :, 0], net[:, :, 1], net[:, :, 2]
    ones = np.ones(S)
    return np.dot(ones, net[:, :, 0]), np.dot(ones, net[:, :, 1]), np.dot(ones, net[:, :, 2])
