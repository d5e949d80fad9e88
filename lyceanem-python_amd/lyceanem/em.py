"""Drop-in for the reference's native module `lyceanem.em`
(lyceanem/src/frequency_domain_models.cu:314-320): tile binning and the
per-pair tile / brute-force scattering calls, on the gfx950 kernels.

Same function names, positional arguments, return shapes and error messages
(RuntimeError on invalid input).  Semantics:
- masks use the native culled predicate (raycasting.cuh:19-68) over every
  triangle of the binned table, cast through the uniform grid (the tile scan of
  raycasting_accelerated.cuh:240-264 visits a superset of the cells the ray
  crosses, so it tests the same triangles up to its float32 tile rounding at
  cell borders, which is not reproduced);
- `not_self_to_self == False` leaves source row == sink row pairs untested
  (raycasting_accelerated.cuh:72);
- the per-pair field is em_wave (em.cuh:64-120): the launch projector of the
  source weight times the lossy-propagation factor with the sink normal; sink
  weights are not used.  Path length and phase are float64 here (the native
  code is float32 throughout).
- calculate_scattering_brute_force returns the computed network (the native
  version never copies its result back and returns zeros,
  raycasting_brute_force.cuh:159-171).
"""
from __future__ import annotations

import numpy as np

from . import _native


def _check2d3(a, name):
    a = np.asarray(a)
    if a.ndim != 2 or a.shape[1] != 3:
        raise RuntimeError(f"{name} must be a 2D array with 3 columns")
    return a


def _check1d(a, name):
    a = np.asarray(a)
    if a.ndim != 1:
        raise RuntimeError(f"{name} must be a 1D array")
    return a


def _tile_bins(tri_vertex, tri_index, ncells_x, ncells_y, ymin, cell_width, z_min):
    """(T, 4) int64 tile bins of each triangle, -1 where absent
    (terrain_acceleration_build.cuh:30-113): the (y, z) cells of its three
    vertices (float32 floor((v - min) / width), as bin_point computes them) in
    vertex order without repeats, plus the cell completing the 2 x 2 block of
    an L-shaped triple.  Vertices outside the bounds, or a triangle whose
    vertex cells are not adjacent (check_triangle_side_length, :114-134, a
    device assert there), raise RuntimeError."""
    v = np.asarray(tri_vertex, np.float32)
    t = np.asarray(tri_index, np.int64).reshape(-1, 3)
    ny = int(ncells_y)
    by = np.floor((v[:, 1] - np.float32(ymin)) / np.float32(cell_width)).astype(np.int64)
    bz = np.floor((v[:, 2] - np.float32(z_min)) / np.float32(cell_width)).astype(np.int64)
    if (by < 0).any() or (by >= ncells_x).any() or (bz < 0).any() or (bz >= ny).any():
        raise RuntimeError("triangle vertices outside the provided y-z bounds")
    b = by * ny + bz
    b0, b1, b2 = b[t[:, 0]], b[t[:, 1]], b[t[:, 2]]
    (r0, c0), (r1, c1), (r2, c2) = divmod(b0, ny), divmod(b1, ny), divmod(b2, ny)
    near = lambda ra, ca, rb, cb: (np.abs(ra - rb) <= 1) & (np.abs(ca - cb) <= 1)  # noqa: E731
    if not (near(r0, c0, r1, c1) & near(r0, c0, r2, c2) & near(r1, c1, r2, c2)).all():
        raise RuntimeError("triangle vertices not in adjacent tiles: a triangle edge is "
                           "longer than one tile")
    out = np.full((len(t), 4), -1, np.int64)
    out[:, 0] = b0
    out[:, 1] = np.where(b1 != b0, b1, -1)
    third = (b2 != b0) & (b2 != b1)
    out[:, 2] = np.where(third, b2, -1)
    # the fourth cell: a coordinate shared by two vertex cells is taken from the
    # third one, pair by pair in the reference's order
    rx = np.full(len(t), -1, np.int64)
    ry = np.full(len(t), -1, np.int64)
    for (ra, ca), (rb, cb), (rc, cc) in (((r0, c0), (r1, c1), (r2, c2)),
                                         ((r0, c0), (r2, c2), (r1, c1)),
                                         ((r1, c1), (r2, c2), (r0, c0))):
        rx = np.where(ra == rb, rc, rx)
        ry = np.where(ca == cb, cc, ry)
    four = third & (b1 != b0)
    out[four, 3] = rx[four] * ny + ry[four]
    return out


def bin_counts_to_numpy(tri_vertex, tri_index, ncells_x, ncells_y, ymin, cell_width, z_min):
    """Triangles per (y, z) tile (terrain_acceleration_build.cuh:296-348)."""
    _check2d3(tri_vertex, "tri_vertex")
    _check2d3(tri_index, "tri_index")
    bins = _tile_bins(tri_vertex, tri_index, ncells_x, ncells_y, ymin, cell_width, z_min)
    return np.bincount(bins[bins >= 0], minlength=ncells_x * ncells_y).astype(np.int32)


def bin_triangles_to_numpy(tri_vertex, tri_index, ncells_x, ncells_y, ymin, cell_width, z_min,
                           bin_counts, sum_bin_counts):
    """Vertex-index triples of the binned triangles, tile by tile
    (terrain_acceleration_build.cuh:370-431); within a tile in triangle order
    (the reference fills tiles with atomics, in no defined order).  Entries
    beyond the binned count stay -1."""
    _check2d3(tri_vertex, "tri_vertex")
    t = _check2d3(tri_index, "tri_index")
    bc = np.asarray(bin_counts)
    if bc.ndim != 1 or bc.shape[0] != ncells_x * ncells_y:
        raise RuntimeError("bincounts is a 1d flattened 2d array thus must have 1d with "
                           "ncountsy *ncountz")
    bins = _tile_bins(tri_vertex, t, ncells_x, ncells_y, ymin, cell_width, z_min)
    tri, slot = np.nonzero(bins >= 0)  # triangle-major, so a stable sort keeps triangle order
    order = np.argsort(bins[tri, slot], kind="stable")
    rows = np.asarray(t, np.int32)[tri[order]]
    out = np.full((int(sum_bin_counts), 3), -1, np.int32)
    n = min(len(out), len(rows))
    out[:n] = rows[:n]
    return out.ravel()


def _pair_network(source, end, tris9, ex, normal, alpha, beta, not_self_to_self):
    S, K = len(source), len(end)
    skip = not bool(not_self_to_self)
    if len(tris9) >= 4096:
        g = _native.Grid(tris9)
        bits, _ = g.visibility(source, end, semantics=_native.SEM_NATIVE, skip_self=skip)
    else:
        bits, _ = _native.visibility_dense(source, end, tris9, semantics=_native.SEM_NATIVE,
                                           any_hit=True, skip_self=skip)
    snk_n = np.asarray(normal, np.float32)[S:S + K]
    net = _native.field_los(source, ex, end, snk_n, hit_bits=bits, alpha=alpha, beta=beta,
                            elements=True)
    return net.cpu().numpy().view(np.float32).reshape(-1)


def _weights(ex_real, ex_imag, ey_real, ey_imag, ez_real, ez_imag, S):
    arrs = [_check1d(a, n) for a, n in ((ex_real, "ex_real"), (ex_imag, "ex_imag"),
                                        (ey_real, "ey_real"), (ey_imag, "ey_imag"),
                                        (ez_real, "ez_real"), (ez_imag, "ez_imag"))]
    for a, n in zip(arrs, ("ex_real", "ex_imag", "ey_real", "ey_imag", "ez_real", "ez_imag")):
        if a.shape[0] != S:
            raise RuntimeError(f"Mismatch: {n}.shape(0) != source_size")
    w = np.empty((S, 3), np.complex64)
    w[:, 0] = arrs[0] + 1j * arrs[1]
    w[:, 1] = arrs[2] + 1j * arrs[3]
    w[:, 2] = arrs[4] + 1j * arrs[5]
    return w


def calculate_scattering_tiles(source, end, triangle_vertex, wave_length, ex_real, ex_imag,
                               ey_real, ey_imag, ez_real, ez_imag, normal, xmin, xmax, ymin, ymax,
                               zmin, zmax, tile_size, bin_count, bin_triangles, n_cellsx,
                               n_cellsy, binned_tri_num, alpha, beta, not_self_to_self):
    """Per-pair network S x K x 3 as interleaved float32 (re, im)
    (frequency_domain_models.cu:44-187)."""
    source = _check2d3(source, "source").astype(np.float32)
    end = _check2d3(end, "end").astype(np.float32)
    tv = _check2d3(triangle_vertex, "triangle_vertex").astype(np.float32)
    nrm = _check2d3(normal, "normal")
    bc = _check1d(bin_count, "bin_count")
    bt = _check1d(bin_triangles, "bin_triangles")
    if bc.shape[0] != n_cellsx * n_cellsy:
        raise RuntimeError("bin_count must have n_cellsx * n_cellsy elements")
    if bt.shape[0] != binned_tri_num * 3:
        raise RuntimeError("bin_triangles 3* bintrinum elements")
    S, K = len(source), len(end)
    w = _weights(ex_real, ex_imag, ey_real, ey_imag, ez_real, ez_imag, S)
    if nrm.shape[0] != S + K:
        raise RuntimeError("Mismatch: normal.shape(0) != points_size")
    tri_idx = np.unique(bt.reshape(-1, 3), axis=0)
    tri_idx = tri_idx[(tri_idx >= 0).all(axis=1)]
    tris9 = tv[tri_idx].reshape(-1, 9)
    return _pair_network(source, end, tris9, w, nrm, alpha, beta, not_self_to_self)


def calculate_scattering_brute_force(source, end, triangles, triangle_vertex, wave_length,
                                     ex_real, ex_imag, ey_real, ey_imag, ez_real, ez_imag, normal,
                                     alpha, beta, not_self_to_self):
    """Per-pair network over every triangle (frequency_domain_models.cu:190-311)."""
    source = _check2d3(source, "source").astype(np.float32)
    end = _check2d3(end, "end").astype(np.float32)
    tri = _check2d3(triangles, "triangles").astype(np.int64)
    tv = _check2d3(triangle_vertex, "triangle_vertex").astype(np.float32)
    nrm = _check2d3(normal, "normal")
    S, K = len(source), len(end)
    w = _weights(ex_real, ex_imag, ey_real, ey_imag, ez_real, ez_imag, S)
    if nrm.shape[0] != S + K:
        raise RuntimeError("Mismatch: normal.shape(0) != points_size")
    tris9 = tv[tri].reshape(-1, 9)
    return _pair_network(source, end, tris9, w, nrm, alpha, beta, not_self_to_self)
