"""Sink generators of the hot path (targets.py subset)."""
from __future__ import annotations

import numpy as np

from ..mesh import Mesh
from .geometryfunctions import compute_areas, theta_phi_r


def spherical_field(az_range, elev_range, outward_normals=False, field_radius=1.0):
    """Points on a sphere at every (elevation, azimuth) pair (targets.py:448-497).

    Point order is elevation-major, azimuth-fastest -- the order the reference's
    consumers reshape to (n_el, n_az), e.g. docs/source/examples/
    02_coherently_polarised_array.py:101-102.  Poles and the +-180 deg seam are
    kept as separate points, as in the reference grid.  Normals point outward
    (outward_normals truthy) or inward; a quad grid split into triangles carries
    cell areas, and theta / phi / radius are attached as point data."""
    az = np.radians(np.asarray(az_range, np.float64))
    el = np.radians(np.asarray(elev_range, np.float64))
    azg, elg = np.meshgrid(az, el)
    r = float(field_radius)
    pts = np.stack([r * np.cos(elg) * np.cos(azg), r * np.cos(elg) * np.sin(azg),
                    r * np.sin(elg)], axis=-1).reshape(-1, 3)
    ne, na = len(el), len(az)
    ii, jj = np.meshgrid(np.arange(ne - 1), np.arange(na - 1), indexing="ij")
    a = (ii * na + jj).ravel()
    b = a + 1
    c = a + na
    d = c + 1
    tris = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)], axis=0)
    mesh = Mesh(points=pts, cells=[("triangle", tris.astype(np.int64))])
    nrm = pts / np.linalg.norm(pts, axis=1).reshape(-1, 1)
    mesh.point_data["Normals"] = nrm if outward_normals else -nrm
    mesh = compute_areas(mesh)
    return theta_phi_r(mesh)
