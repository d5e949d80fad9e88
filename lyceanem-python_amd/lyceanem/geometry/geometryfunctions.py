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


def compute_areas(field_data):
    """Triangle areas as cell data and one third of each adjacent triangle's area
    per point as point data."""
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
    field_data.cell_data["Area"] = [a]
    field_data.point_data["Area"] = pa
    return field_data


def compute_normals(mesh):
    """Unit triangle normals (cell data) and area-weighted unit point normals."""
    tri = triangle_cells(mesh)
    p = np.asarray(mesh.points, np.float64)
    n = np.cross(p[tri[:, 1]] - p[tri[:, 0]], p[tri[:, 2]] - p[tri[:, 0]])
    pn = np.zeros_like(p)
    for j in range(3):
        np.add.at(pn, tri[:, j], n)
    with np.errstate(invalid="ignore", divide="ignore"):
        cn = n / np.linalg.norm(n, axis=1, keepdims=True)
        pn = pn / np.linalg.norm(pn, axis=1, keepdims=True)
    mesh.cell_data["Normals"] = [cn]
    mesh.point_data["Normals"] = pn
    return mesh


def cell_centroids(field_data):
    """Triangle centroids as a point cloud carrying the cell data as point data."""
    tri = triangle_cells(field_data)
    p = np.asarray(field_data.points, np.float64)
    c = (1.0 / 3.0) * (p[tri[:, 0]] + p[tri[:, 1]] + p[tri[:, 2]])
    pd = {k: np.asarray(v[-1]) for k, v in field_data.cell_data.items()}
    return Mesh(points=c, cells=[("vertex", np.arange(len(c)).reshape(-1, 1))], point_data=pd)
