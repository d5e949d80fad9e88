"""Geometry helpers of the hot path (geometryfunctions.py subset)."""
from __future__ import annotations

import numpy as np

from .. import base_types
from ..mesh import Mesh, triangle_cells


def _as_meshes(obj):
    """Flatten the solids a caller can pass to mesh_conversion."""
    if obj is None:
        return []
    if isinstance(obj, (list, tuple)):
        out = []
        for item in obj:
            out.extend(_as_meshes(item))
        return out
    if hasattr(obj, "export_all_structures"):
        return _as_meshes(obj.export_all_structures())
    if hasattr(obj, "solids"):
        return _as_meshes(list(obj.solids))
    if hasattr(obj, "points") and hasattr(obj, "cells"):
        return [obj]
    return []


def mesh_conversion(conversion_object):
    """Solids -> triangle_t records (geometryfunctions.py:473-510).

    Accepts a mesh (meshio.Mesh or lyceanem.mesh.Mesh), a list of meshes, or an
    object with `solids` / `export_all_structures()`; triangles are appended in
    that order."""
    from ..raycasting.rayfunctions import convertTriangles

    parts = [convertTriangles(m) for m in _as_meshes(conversion_object)]
    if not parts:
        return np.empty(0, dtype=base_types.triangle_t)
    return np.concatenate(parts)


def theta_phi_r(field_data):
    """Spherical coordinates of the points about the origin
    (geometryfunctions.py:447-470)."""
    p = np.asarray(field_data.points, dtype=np.float64)
    r = np.linalg.norm(p - np.zeros((1, 3)), axis=1)
    field_data.point_data["Radial_Distance_(m)"] = r
    field_data.point_data["theta_(Radians)"] = np.arccos(p[:, 2] / r)
    field_data.point_data["phi_(Radians)"] = np.arctan2(p[:, 1], p[:, 0])
    return field_data


def _copy_mesh(mesh):
    """A new mesh with the same points, cells and data (the reference's area and
    normal helpers return the pyvista round trip's new mesh, leaving theirs
    unchanged)."""
    return Mesh(points=np.array(mesh.points, copy=True), cells=list(mesh.cells),
                point_data=dict(mesh.point_data), cell_data=dict(mesh.cell_data))


def locate_cell_index(field_data, cell_type="triangle"):
    """Index of the last cell block of cell_type (geometryfunctions.py:228-252:
    the loop keeps overwriting the index, so the last block wins)."""
    desired = None
    for inc, cell in enumerate(field_data.cells):
        if cell.type == cell_type:
            desired = inc
    if desired is None:
        raise ValueError(f"no {cell_type} cells in the mesh")
    return desired


def compute_areas(field_data):
    """A copy of the mesh with triangle areas as cell data and one third of
    each adjacent triangle's area per point as point data
    (geometryfunctions.py:255-356)."""
    field_data = _copy_mesh(field_data)
    tri = triangle_cells(field_data)
    p = np.asarray(field_data.points, np.float64)
    if len(tri):
        a = 0.5 * np.linalg.norm(
            np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]]), axis=1)
    else:
        a = np.zeros(0)
    pa = np.zeros(len(p))
    for j in range(3):
        np.add.at(pa, tri[:, j], a / 3.0)
    field_data.cell_data["Area"] = _per_block(field_data, a)
    field_data.point_data["Area"] = pa
    return field_data


def _per_block(mesh, values):
    """Cell data list: `values` for the triangle block, empty arrays for others."""
    idx = locate_cell_index(mesh) if any(c.type == "triangle" for c in mesh.cells) else None
    if idx is None:
        return [values]
    return [values if i == idx else np.zeros((len(c.data),) + values.shape[1:])
            for i, c in enumerate(mesh.cells)]


def compute_normals(mesh):
    """A copy of the mesh with unit triangle normals (cell data) and
    area-weighted unit point normals (geometryfunctions.py:359-444)."""
    mesh = _copy_mesh(mesh)
    tri = triangle_cells(mesh)
    p = np.asarray(mesh.points, np.float64)
    n = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]])
    pn = np.zeros_like(p)
    for j in range(3):
        np.add.at(pn, tri[:, j], n)
    with np.errstate(invalid="ignore", divide="ignore"):
        cn = n / np.linalg.norm(n, axis=1, keepdims=True)
        pn = pn / np.linalg.norm(pn, axis=1, keepdims=True)
    mesh.cell_data["Normals"] = _per_block(mesh, cn)
    mesh.point_data["Normals"] = pn
    return mesh


def cell_centroids(field_data):
    """Triangle centroids as a point cloud carrying the triangle block's cell
    data as point data (geometryfunctions.py:45-92)."""
    tri = triangle_cells(field_data)
    p = np.asarray(field_data.points, np.float64)
    c = (1.0 / 3.0) * (p[tri[:, 0]] + p[tri[:, 1]] + p[tri[:, 2]])
    idx = locate_cell_index(field_data)
    pd = {k: np.asarray(v[idx]) for k, v in field_data.cell_data.items()}
    return Mesh(points=c, cells=[("vertex", np.arange(len(c)).reshape(-1, 1))], point_data=pd)
