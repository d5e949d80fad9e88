"""numpy record layouts shared with the reference (base_types.py:21-119).

The kernels and the reference's own host code exchange these records, so the
field names, order, types and alignment are kept identical; the numba type
objects the reference derives from them are not needed here.
"""
import numpy as np

point_data = np.dtype(
    [("permittivity", "c8"), ("permeability", "c8"), ("Electric", "?")], align=True
)

# v0, v1, v2 as 9 float32 (36 bytes) -- read by the kernels as a T x 9 array.
triangle_t = np.dtype(
    [(f"v{i}{c}", "f4") for i in range(3) for c in "xyz"], align=True
)

# origin, unit direction, length, blocked flag (32 bytes with alignment).
ray_t = np.dtype(
    [("ox", "f4"), ("oy", "f4"), ("oz", "f4"), ("dx", "f4"), ("dy", "f4"), ("dz", "f4"),
     ("dist", "f4"), ("intersect", "?")],
    align=True,
)

scattering_t = np.dtype(
    [("px", "f4"), ("py", "f4"), ("pz", "f4"),
     ("vx", "f4"), ("vy", "f4"), ("vz", "f4"),
     ("nx", "f4"), ("ny", "f4"), ("nz", "f4"),
     ("ex", "c8"), ("ey", "c8"), ("ez", "c8"),
     ("permittivity", "c8"), ("permeability", "c8"), ("Electric", "?")],
    align=True,
)
scattering_point = scattering_t


def triangles_as_array(triangles) -> np.ndarray:
    """triangle_t records (or any (T,9)/(T,3,3) float array) -> (T, 9) float32."""
    a = np.asarray(triangles)
    if a.dtype.names is not None:
        if a.dtype != triangle_t:
            a = a.astype(triangle_t)
        return np.ascontiguousarray(a).view(np.float32).reshape(-1, 9)
    return np.ascontiguousarray(a, dtype=np.float32).reshape(-1, 9)
