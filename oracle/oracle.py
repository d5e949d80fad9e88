"""CPU oracle for LyceanEM's raycast + frequency-domain hot path.

TEST INFRASTRUCTURE ONLY.  Imported by tests/, by ``__graft_entry__.smoke()`` and
by the ``cpu_baseline`` leg of ``bench.py`` -- always as the checker (or the timed
CPU baseline), never as the product.  The product package never imports it.

Two parts:

* ``liblem_oracle.so`` (``lem_oracle.c``, built by ``oracle/Makefile``): the
  float32 Moller-Trumbore predicates of the reference (numba two-sided,
  rayfunctions.py:37-104, and native culled, raycasting.cuh:19-68), the
  kernel1Dv3 closest-hit loop (rayfunctions.py:212-244) and the float64 field
  restatement of freqdomainkernal / lossy_propagation / sourcelaunchtransformGPU
  (empropagation.py:178-249, 768-801, 1083-1265).
* Pure-Python/numpy restatements of the host index semantics:
  create_model_index (rayfunctions.py:2624-2678), rayHits1Dv2 (:1845-1878),
  workchunkingv2's wave assembly (:2681-2955), targettingindex
  (empropagation.py:3394-3429), theta_phi_r (geometryfunctions.py:447-470) and
  the E-theta / E-phi conversion of calculate_farfield (frequency_domain.py:179-190).
* The time-domain per-path amplitude / arrival and placement of timedomainkernal
  (empropagation.py:1345-1612; C: orc_time_paths, orc_time_place), the
  projected-area accumulation of visiblespace (rayfunctions.py:316-427) and the
  directivity map of MaximumDirectivityMap / directivity_transform
  (beamforming.py:484-673, 1037-1129).

Every function is pinned by a check in tests/test_oracle_*.py that does not
restate the same formula (closed forms, exact rational geometry, invariants,
Python's own floor division, literal loops of the reference's cell sums).
Parity unpinned, because no output of the reference fixes them and the
reference cannot run here: numba-cuda's possible FMA contraction inside `hit`
and `dot_vec` (everything here is evaluated without contraction), numba's
float32 typing of calc_dv_norm (assumed float32), the order of the reference's
float64 atomic sums, and pyvista's sink order, areas and normals.
"""
from __future__ import annotations

import ctypes
import os
import subprocess
from fractions import Fraction

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "liblem_oracle.so")
_lib = None


def build(force: bool = False) -> str:
    """Compile liblem_oracle.so with the committed Makefile."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-s", "-B" if force else "all"], cwd=_HERE, check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        P = ctypes.c_void_p
        i64 = ctypes.c_int64
        L.orc_hit_numba.argtypes = [P, P, ctypes.c_float, P, P]
        L.orc_hit_numba.restype = ctypes.c_int
        L.orc_hit_native.argtypes = [P, P, ctypes.c_float, P]
        L.orc_hit_native.restype = ctypes.c_int
        L.orc_charge.argtypes = [P, P, i64, P, P]
        L.orc_raycast_rays.argtypes = [P, i64, P, i64, P, P, ctypes.c_int]
        L.orc_visibility_dense.argtypes = [P, i64, P, i64, P, i64, ctypes.c_int,
                                           ctypes.c_int, ctypes.c_int, P, P, ctypes.c_int]
        L.orc_visibility_sampled.argtypes = [P, P, P, i64, P, i64, ctypes.c_int, P, P,
                                             ctypes.c_int]
        L.orc_field_los.argtypes = [P, P, i64, P, P, P, i64, P, ctypes.c_double,
                                    ctypes.c_double, ctypes.c_int, P, ctypes.c_int]
        L.orc_field_paths.argtypes = [P, P, P, P, i64, P, i64, ctypes.c_int, i64, i64,
                                      ctypes.c_double, ctypes.c_double, ctypes.c_int, P]
        f32 = ctypes.c_float
        L.orc_tiles_build.argtypes = [P, i64, ctypes.c_int, f32, f32, f32, ctypes.c_int,
                                      ctypes.c_int, f32, f32]
        L.orc_tiles_build.restype = P
        L.orc_tiles_refs.argtypes = [P]
        L.orc_tiles_refs.restype = i64
        L.orc_tiles_free.argtypes = [P]
        L.orc_tiles_cast.argtypes = [P, P, P, P, i64, ctypes.c_int, P, ctypes.c_int]
        L.orc_py_floordiv.argtypes = [ctypes.c_double, ctypes.c_double]
        L.orc_py_floordiv.restype = ctypes.c_double
        L.orc_time_paths.argtypes = [P, P, P, P, i64, P, i64, ctypes.c_int, ctypes.c_double,
                                     ctypes.c_double, P, P]
        L.orc_time_place.argtypes = [P, P, i64, P, P, i64, ctypes.c_double, i64, P, P, P]
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def nthreads_default() -> int:
    try:
        return len(os.sched_getaffinity(0))
    except AttributeError:  # pragma: no cover
        return os.cpu_count() or 1


# --------------------------------------------------------------------------
# predicates and raycasting
# --------------------------------------------------------------------------

def hit_numba(o, d, dist, tri9):
    """numba `hit` (rayfunctions.py:37-104): returns (bool, float64 t or inf)."""
    o, d, tri = _f32(o), _f32(d), _f32(tri9)
    t = np.zeros(1, np.float64)
    h = lib().orc_hit_numba(_p(o), _p(d), ctypes.c_float(dist), _p(tri), _p(t))
    return bool(h), (float(t[0]) if h else float("inf"))


def hit_native(o, d, length, tri9):
    """native CUDA `hit` (raycasting.cuh:19-68): culled, eps = 1e-6f."""
    o, d, tri = _f32(o), _f32(d), _f32(tri9)
    return bool(lib().orc_hit_native(_p(o), _p(d), ctypes.c_float(length), _p(tri)))


def charge(origins, targets):
    """calc_dv_norm (math_functions.py:148-156): float32 unit directions, lengths."""
    o, t = _f32(origins).reshape(-1, 3), _f32(targets).reshape(-1, 3)
    n = o.shape[0]
    d = np.empty((n, 3), np.float32)
    L = np.empty(n, np.float32)
    lib().orc_charge(_p(o), _p(t), n, _p(d), _p(L))
    return d, L


def raycast_rays(rays7, tris9, nthreads=None):
    """kernel1Dv3 over a charged payload; rays7 is (R,7) (ox..dz, dist)."""
    r = _f32(rays7).reshape(-1, 7)
    tr = _f32(tris9).reshape(-1, 9)
    R = r.shape[0]
    inter = np.zeros(R, np.uint8)
    dist = np.zeros(R, np.float32)
    lib().orc_raycast_rays(_p(r), R, _p(tr), tr.shape[0], _p(inter), _p(dist),
                           nthreads or nthreads_default())
    return inter.astype(bool), dist


def visibility_dense(src, snk, tris9, semantics=0, any_hit_only=False, skip_self=False,
                     want_dist=True, nthreads=None):
    """Dense S x K hit mask (True = blocked) and kernel1Dv3 `dist`."""
    s, k = _f32(src).reshape(-1, 3), _f32(snk).reshape(-1, 3)
    tr = _f32(tris9).reshape(-1, 9)
    S, K = s.shape[0], k.shape[0]
    hit = np.zeros((S, K), np.uint8)
    dist = np.zeros((S, K), np.float32) if want_dist else None
    lib().orc_visibility_dense(_p(s), S, _p(k), K, _p(tr), tr.shape[0], int(semantics),
                               int(any_hit_only), int(skip_self), _p(hit), _p(dist),
                               nthreads or nthreads_default())
    return hit.astype(bool), dist


def visibility_sampled(src, snk, pairs, tris9, any_hit_only=False, nthreads=None):
    """numba semantics on explicit (source row, sink row) pairs."""
    s, k = _f32(src).reshape(-1, 3), _f32(snk).reshape(-1, 3)
    pr = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
    tr = _f32(tris9).reshape(-1, 9)
    n = pr.shape[0]
    hit = np.zeros(n, np.uint8)
    dist = np.zeros(n, np.float32)
    lib().orc_visibility_sampled(_p(s), _p(k), _p(pr), n, _p(tr), tr.shape[0],
                                 int(any_hit_only), _p(hit), _p(dist),
                                 nthreads or nthreads_default())
    return hit.astype(bool), dist


class Tiles:
    """The reference's tile acceleration (Tile_acceleration_structure,
    models/acceleration_structures.py:57-70, with the binning of
    terrain_acceleration_build.cuh:57-133 and the tile DDA of
    raycasting_accelerated.cuh:106-268), restated in C: triangles binned into
    2-D tiles over the two axes other than `up` by their vertices, and for each
    ray every tile of the rectangle between where it enters and leaves the
    structure's slab scanned until a triangle hits.  Bounds carry the
    reference's 0.1 % padding; the tile is min(horizontal extents) / n_cells.
    Any-hit masks only (the reference's structure returns no distances)."""

    def __init__(self, tris9, up=2, n_cells=256):
        self.tr = _f32(tris9).reshape(-1, 9)
        pts = self.tr.reshape(-1, 3).astype(np.float64)
        mx, mn = pts.max(0), pts.min(0)
        hi, lo = mx + np.abs(mx) * 0.001, mn - np.abs(mn) * 0.001
        a, b = [x for x in range(3) if x != up]
        tile = min(hi[a] - lo[a], hi[b] - lo[b]) / n_cells
        self.na = int(np.ceil((hi[a] - lo[a]) / tile))
        self.nb = int(np.ceil((hi[b] - lo[b]) / tile))
        self.tile = tile
        self.h = lib().orc_tiles_build(_p(self.tr), self.tr.shape[0], int(up), float(lo[a]),
                                       float(lo[b]), float(tile), self.na, self.nb,
                                       float(lo[up]), float(hi[up]))
        self.refs = int(lib().orc_tiles_refs(self.h))

    def cast(self, src, snk, pairs, sem=0, nthreads=None):
        """Blocked flags of (source row, sink row) pairs (sem 0: numba
        predicate, 1: native culled)."""
        s, k = _f32(src).reshape(-1, 3), _f32(snk).reshape(-1, 3)
        pr = np.ascontiguousarray(pairs, dtype=np.int64).reshape(-1, 2)
        hit = np.zeros(pr.shape[0], np.uint8)
        lib().orc_tiles_cast(self.h, _p(s), _p(k), _p(pr), pr.shape[0], int(sem), _p(hit),
                             nthreads or nthreads_default())
        return hit.astype(bool)

    def __del__(self):
        if getattr(self, "h", None):
            lib().orc_tiles_free(self.h)
            self.h = None


# --------------------------------------------------------------------------
# fields
# --------------------------------------------------------------------------

def _c64x3(w, n):
    w = np.asarray(w)
    if w.ndim == 1:
        w = np.broadcast_to(w.reshape(1, 3), (n, 3))
    return np.ascontiguousarray(w, dtype=np.complex64).view(np.float32).reshape(n, 6)


def field_los(src_pos, src_w, snk_pos, snk_nrm, alpha, beta, hit=None, snk_w=None,
              sink_tangent=False, elements=False, r_f32=False, nthreads=None):
    """Float64 line-of-sight field (freqdomainkernal per path).

    Returns (K,3) complex128 summed over sources, or (S,K,3) with elements=True.
    """
    sp = _f32(src_pos).reshape(-1, 3)
    S = sp.shape[0]
    kp = _f32(snk_pos).reshape(-1, 3)
    kn = _f32(snk_nrm).reshape(-1, 3)
    K = kp.shape[0]
    sw = _c64x3(src_w, S)
    kw = None if snk_w is None else _c64x3(snk_w, K)
    h = None if hit is None else np.ascontiguousarray(hit, dtype=np.uint8).reshape(S, K)
    flags = (1 if sink_tangent else 0) | (2 if elements else 0) | (4 if r_f32 else 0)
    out = np.zeros((S, K, 6) if elements else (K, 6), np.float64)
    lib().orc_field_los(_p(sp), _p(sw), S, _p(kp), _p(kn), _p(kw), K, _p(h), float(alpha),
                        float(beta), flags, _p(out), nthreads or nthreads_default())
    return out.view(np.complex128).reshape(((S, K, 3) if elements else (K, 3)))


def _eta(eta, N):
    return None if eta is None else np.ascontiguousarray(np.broadcast_to(eta, (N,)), np.float64)


def field_paths(pos, nrm, w, index, S, K, alpha, beta, r_f32=False, eta=None):
    """freqdomainkernal over an explicit 1-based path index -> (S,K,3) complex128.
    eta: per-point impedance, > 0 for magnetic points (None: all electric)."""
    p = _f32(pos).reshape(-1, 3)
    n = _f32(nrm).reshape(-1, 3)
    N = p.shape[0]
    ww = _c64x3(w, N)
    et = _eta(eta, N)
    idx = np.ascontiguousarray(index, dtype=np.int64)
    V, ncol = idx.shape
    out = np.zeros((S, K, 6), np.float64)
    lib().orc_field_paths(_p(p), _p(n), _p(ww), _p(et), N, _p(idx), V, ncol, S, K, float(alpha),
                          float(beta), int(r_f32), _p(out))
    return out.view(np.complex128).reshape(S, K, 3)


# --------------------------------------------------------------------------
# host index semantics (pure Python restatements)
# --------------------------------------------------------------------------

def create_model_index(source_index, sink_index, scattering_point_index):
    """rayfunctions.py:2624-2678: for each source row, one row per target
    (sinks then scattering points), the source row's columns followed by the
    target; rows whose last two columns are equal are dropped."""
    rows = []
    src = np.asarray(source_index)
    if src.ndim == 1:
        src = src.reshape(-1, 1)
    targets = [int(x) for x in np.asarray(sink_index).ravel()]
    targets += [int(x) for x in np.asarray(scattering_point_index).ravel()]
    for n in range(src.shape[0]):
        for t in targets:
            row = [int(v) for v in src[n]] + [t]
            if row[-2] != row[-1]:
                rows.append(row)
    return np.asarray(rows, dtype=np.int32).reshape(-1, src.shape[1] + 1)


def ray_hits(intersect, point_indexing, sink_index):
    """rayHits1Dv2 (rayfunctions.py:1845-1878): (filtered, final) index rows."""
    sinks = set(int(x) for x in np.asarray(sink_index).ravel())
    fin, fil = [], []
    for row, inter in zip(np.asarray(point_indexing), np.asarray(intersect)):
        if inter:
            continue
        (fin if int(row[-1]) in sinks else fil).append(row)
    ncol = np.asarray(point_indexing).shape[1]
    return (np.asarray(fil, np.int32).reshape(-1, ncol),
            np.asarray(fin, np.int32).reshape(-1, ncol))


def workchunking(src, snk, scat, tris9, max_scatter, line_of_sight=True):
    """workchunkingv2 (rayfunctions.py:2681-2955) for max_scatter 1, 2 and 3
    without memory chunking: returns (full_index int32, RAYS_CAST)."""
    src, snk = _f32(src).reshape(-1, 3), _f32(snk).reshape(-1, 3)
    scat = _f32(scat).reshape(-1, 3)
    S, K, P = len(src), len(snk), len(scat)
    unified = np.concatenate([src, snk, scat], axis=0)
    source_index = np.arange(1, S + 1).reshape(S, 1)
    sink_index = np.arange(S + 1, S + 1 + K).reshape(K, 1)
    scat_index = np.arange(S + K + 1, S + K + 1 + P).reshape(P, 1)

    def launch(index):
        o = unified[index[:, -2] - 1]
        t = unified[index[:, -1] - 1]
        d, L = charge(o, t)
        rays = np.concatenate([o, d, L.reshape(-1, 1)], axis=1)
        inter, _ = raycast_rays(rays, tris9)
        fil, fin = ray_hits(inter, index, sink_index)
        return fil, fin, index.shape[0]

    if max_scatter == 1:
        io = create_model_index(source_index, sink_index, scat_index)
        _, fin, n = launch(io)
        return fin.reshape(-1, 2).astype(np.int32), n
    if max_scatter == 3:
        # rayfunctions.py:2863-2953: sources in chunks (np.linspace bounds);
        # per chunk wave 1 (sinks and scatterers), wave 2 from the surviving
        # scatterer rows to sinks and scatterers (chunkingRaycaster1Dv3 with
        # terminate_flag False, :2120-2166), wave 3 from the rows that reached a
        # second scatterer to the sinks; rows padded with zero columns to 4
        estimate = S * (K + P) + S * P * K * 3
        chunknum = max(2, min(S, int(np.ceil(estimate / 2e8))))
        bounds = np.linspace(0, S, chunknum, dtype=np.int32)
        parts, rays = [], 0
        for c in range(len(bounds) - 1):
            srows = source_index[bounds[c]:bounds[c + 1]]
            io = create_model_index(srows, sink_index if line_of_sight else np.empty((0, 1)),
                                    scat_index)
            fil, fin, n1 = launch(io)
            io2 = create_model_index(fil, sink_index, scat_index)
            fil2, fin2, n2 = launch(io2)
            io3 = create_model_index(fil2, sink_index, np.empty((0, 0), np.int32))
            _, fin3, n3 = launch(io3)
            rays = n1 + n2 + n3  # the reference keeps only the last chunk's count
            t = np.concatenate([np.concatenate([fin.reshape(-1, 2),
                                                np.zeros((len(fin), 1), np.int32)], 1),
                                fin2.reshape(-1, 3)])
            parts.append(np.concatenate([np.concatenate([t, np.zeros((len(t), 1), np.int32)], 1),
                                         fin3.reshape(-1, 4)]))
        return np.concatenate(parts).astype(np.int32).reshape(-1, 4), rays
    if max_scatter != 2:
        raise NotImplementedError("oracle covers max_scatter 1, 2 and 3")
    if line_of_sight:
        io = create_model_index(source_index, sink_index, scat_index)
    else:
        io = create_model_index(source_index, np.empty((0, 1)), scat_index)
    fil, fin, n1 = launch(io)
    if fil.shape[0] == 0:
        # rayfunctions.py:2835-2861: the line-of-sight rows are not appended
        # when no scatterer is visible (the append sits in the else branch).
        return np.empty((0, 3), np.int32), n1
    io2 = create_model_index(fil, sink_index, np.empty((0, 0), np.int32))
    _, fin2, n2 = launch(io2)
    full = np.concatenate([np.concatenate([fin, np.zeros((fin.shape[0], 1), np.int32)], 1),
                           fin2], axis=0)
    return full.astype(np.int32), n1 + n2


def targetting_index(full_index):
    """targettingindex (empropagation.py:3394-3429) for 2 and 3 columns."""
    fi = np.asarray(full_index)
    if fi.shape[1] == 2:
        return fi.copy(), np.ones(fi.shape[0], np.int64)
    los = [r[[0, 1]] for r in fi if r[2] == 0]
    sc = [r[[0, 2]] for r in fi if r[2] != 0]
    depth = np.asarray(los + sc, dtype=fi.dtype).reshape(-1, 2)
    scatter = np.where(fi[:, 2] != 0, 2, 1).astype(np.int64)
    return depth, scatter


def theta_phi(points):
    """theta_phi_r (geometryfunctions.py:447-470)."""
    p = np.asarray(points, np.float64)
    r = np.linalg.norm(p, axis=1)
    return np.arccos(p[:, 2] / r), np.arctan2(p[:, 1], p[:, 0]), r


def etheta_ephi(Ex, Ey, Ez, theta, phi):
    """calculate_farfield's conversion (frequency_domain.py:179-190)."""
    et = (Ex * np.cos(phi) * np.cos(theta) + Ey * np.sin(phi) * np.cos(theta)
          - Ez * np.sin(theta))
    ep = -Ex * np.sin(phi) + Ey * np.cos(phi)
    return et, ep


# --------------------------------------------------------------------------
# exact geometry (pins the predicates away from their rounding band)
# --------------------------------------------------------------------------

def exact_segment_triangle(o, t, tri9):
    """Exact rational intersection of segment o->t with a triangle.

    Returns (inside_margin, s) where s in [0,1] is the segment parameter of the
    plane crossing, and inside_margin the smallest barycentric coordinate (the
    signed distance to the nearest edge in barycentric units), or None when the
    segment is parallel to the plane."""
    O = [Fraction(float(x)) for x in o]
    T = [Fraction(float(x)) for x in t]
    V = [Fraction(float(x)) for x in tri9]
    v0, v1, v2 = V[0:3], V[3:6], V[6:9]
    e1 = [v1[i] - v0[i] for i in range(3)]
    e2 = [v2[i] - v0[i] for i in range(3)]
    D = [T[i] - O[i] for i in range(3)]

    def cross(a, b):
        return [a[1] * b[2] - a[2] * b[1], a[2] * b[0] - a[0] * b[2], a[0] * b[1] - a[1] * b[0]]

    def dot(a, b):
        return a[0] * b[0] + a[1] * b[1] + a[2] * b[2]

    p = cross(D, e2)
    A = dot(e1, p)
    if A == 0:
        return None
    tv = [O[i] - v0[i] for i in range(3)]
    u = dot(tv, p) / A
    q = cross(tv, e1)
    v = dot(D, q) / A
    s = dot(e2, q) / A
    return min(u, v, 1 - u - v), s


# --------------------------------------------------------------------------
# float64 torch restatement of the line-of-sight field (dense reference for
# gradients: autograd of this closed form, and finite differences of it)
# --------------------------------------------------------------------------

def field_los_torch64(w, pos, snk, snk_n, visible, alpha, beta):
    """E_k = sum_s visible[s,k] F(r) (d.n_k) (I - d d^T) w_s in float64 / complex128
    (lossy_propagation, empropagation.py:768-801, with the launch projector
    of :178-249 in closed form).  w: (S,3) complex128, pos: (S,3) float64,
    snk / snk_n: (K,3) float64, visible: (S,K) bool.  Differentiable in w, pos."""
    import torch

    D = snk[None, :, :] - pos[:, None, :]
    r = torch.sqrt((D * D).sum(-1))
    d = D / r[..., None]
    cos = (d * snk_n[None]).sum(-1)
    g = complex(alpha, beta)
    G = torch.exp(-g * r) / r
    F = -(1.0 / (2.0 * np.pi)) * (-g - 1.0 / r) * G
    L = F * cos * visible.to(torch.float64)
    dw = (d.to(torch.complex128) * w[:, None, :]).sum(-1)
    Q = w[:, None, :] - d.to(torch.complex128) * dw[..., None]
    return (L[..., None] * Q).sum(0)


# ---------------------------------------------------------------------------
# time domain (timedomainkernal, empropagation.py:1345-1612)
# ---------------------------------------------------------------------------

def py_floordiv(a, b):
    """Python's float a // b as the C restatement computes it."""
    return lib().orc_py_floordiv(float(a), float(b))


def time_paths(pos, nrm, w, index, alpha, beta, eta=None):
    """(amp (V, 3), arrival (V,)) float64 of every row of a 1-based index;
    eta as for field_paths."""
    p = _f32(pos).reshape(-1, 3)
    n = _f32(nrm).reshape(-1, 3)
    N = p.shape[0]
    ww = _c64x3(w, N)
    et = _eta(eta, N)
    idx = np.ascontiguousarray(index, np.int32)
    V, ncol = idx.shape
    amp = np.zeros((V, 3), np.float64)
    arr = np.zeros(V, np.float64)
    lib().orc_time_paths(_p(p), _p(n), _p(ww), _p(et), N, _p(idx), V, ncol, float(alpha),
                         float(beta), _p(amp), _p(arr))
    return amp, arr


def time_place(amp, arrival, slot, nslots, excitation, fs, N):
    """(out (nslots, N, 3), wake, bins): the reference's per-path placement."""
    amp = np.ascontiguousarray(amp, np.float64)
    arr = np.ascontiguousarray(arrival, np.float64)
    sl = np.ascontiguousarray(slot, np.int64)
    exc = np.ascontiguousarray(excitation, np.float64)
    out = np.zeros((int(nslots), int(N), 3), np.float64)
    wake = np.zeros(1, np.float64)
    bins = np.zeros(len(arr), np.int64)
    lib().orc_time_place(_p(amp), _p(arr), len(arr), _p(sl), _p(exc), len(exc), float(fs),
                         int(N), _p(out), _p(wake), _p(bins))
    return out, float(wake[0]), bins


# ---------------------------------------------------------------------------
# visiblespace projected-area pattern (rayfunctions.py:316-468)
# ---------------------------------------------------------------------------

def _signed_angle(v1, v2):
    """angle() (rayfunctions.py:431-448) row-wise in float32: unit vectors, the
    sign from the 2x2 minor of the last two components, arccos of the clamped
    dot product."""
    v1 = np.asarray(v1, np.float32)
    v2 = np.asarray(v2, np.float32)
    u1 = v1 / np.linalg.norm(v1, axis=1, keepdims=True)
    u2 = v2 / np.linalg.norm(v2, axis=1, keepdims=True)
    minor = u1[:, 1] * u2[:, 2] - u1[:, 2] * u2[:, 1]
    sign = np.where(minor == 0, np.float32(1), -np.sign(minor)).astype(np.float32)
    dp = np.clip(np.sum(u1 * u2, axis=1), -1.0, 1.0).astype(np.float32)
    return (sign * np.arccos(dp)).astype(np.float32)


def projected_area(src, nrm, vertex_area, snk, hit):
    """visiblespace's pattern (rayfunctions.py:366-421): for every unblocked
    (source, shell point) ray, the portion vertex_area * |cos(angle)| (1 when a
    single vertex_area of 0 is given), zero beyond 90 degrees, summed per shell
    point (patternsort).  hit: (S, K) bool, True = blocked."""
    src = _f32(src).reshape(-1, 3)
    snk = _f32(snk).reshape(-1, 3)
    s_idx, k_idx = np.nonzero(~np.asarray(hit, bool))
    d, _ = charge(src[s_idx], snk[k_idx])
    ang = _signed_angle(d, np.asarray(nrm, np.float32)[s_idx])
    va = np.atleast_1d(np.asarray(vertex_area, np.float64)).ravel()
    va = np.where(np.isnan(va), 0.0, va)
    if va.size == 1:
        if va[0] == 0:
            portion = np.ones(len(ang))
        else:
            portion = np.maximum(va[0] * np.abs(np.cos(ang)), 0)
    else:
        portion = va[s_idx] * np.abs(np.cos(ang))
    portion[np.abs(ang) > np.pi / 2] = 0.0
    out = np.zeros(len(snk), np.float64)
    np.add.at(out, k_idx, portion)
    return out.astype(np.float32)


# ---------------------------------------------------------------------------
# beamforming (electromagnetics/beamforming.py)
# ---------------------------------------------------------------------------

def directivity_transform(Etheta, Ephi, az_range, elev_range, total_solid_angle=4 * np.pi):
    """directivity_transform (beamforming.py:1037-1129) as its cell loop:
    power_sum += mean(4 corner Utheta) |sin(rad(az_j + az_j+1) / 2)| + the same
    for Uphi; Uav = power_sum * d_az * d_el / total_solid_angle (radians)."""
    Ut = np.abs(np.asarray(Etheta, np.complex128) ** 2)
    Up = np.abs(np.asarray(Ephi, np.complex128) ** 2)
    power_sum = 0.0
    for i in range(len(elev_range) - 1):
        for j in range(len(az_range) - 1):
            s = abs(np.sin(np.radians(az_range[j] + az_range[j + 1]) / 2.0))
            r1 = np.mean([Ut[i, j], Ut[i + 1, j], Ut[i, j + 1], Ut[i + 1, j + 1]])
            r2 = np.mean([Up[i, j], Up[i + 1, j], Up[i, j + 1], Up[i + 1, j + 1]])
            power_sum += r1 * s + r2 * s
    Uav = power_sum * (np.radians(az_range[1] - az_range[0]) *
                       np.radians(elev_range[1] - elev_range[0])) / total_solid_angle
    return Ut / Uav, Up / Uav, (Ut + Up) / Uav


def max_directivity_map(Etheta, Ephi, points, wavelength, az_range, elev_range,
                        forming="Total", total_solid_angle=4 * np.pi):
    """MaximumDirectivityMap (beamforming.py:484-673) with every grid point a
    command angle, in complex128: per command, equal-gain weights on Etheta and
    on Ephi and wavefront weights, each beamformed pattern's directivity (the
    first with the default 4 pi solid angle), the best by `forming` read at
    (elev_inc, az_inc)."""
    Et = np.asarray(Etheta, np.complex64)
    Ep = np.asarray(Ephi, np.complex64)
    pts = np.asarray(points, np.float64)
    c = 299792458.0
    out = np.zeros((len(elev_range), len(az_range), 3))
    for a, az in enumerate(az_range):
        for e, el in enumerate(elev_range):
            # command_angles is a float32 array (beamforming.py:557): so are the
            # radians and the steering vector sph2cart builds from them
            ra, re_ = np.radians(np.float32(az)), np.radians(np.float32(el))
            sv = np.array([np.cos(re_) * np.cos(ra), np.cos(re_) * np.sin(ra), np.sin(re_)],
                          np.float64)
            dist = np.sqrt(np.sum((pts - sv * 1e9) ** 2, axis=1))
            dist = dist - dist.min()
            ws = np.exp(1j * 2 * np.pi * (c / wavelength) * (dist / c))
            sets = [np.exp(-1j * np.angle(Et[:, e, a])), np.exp(-1j * np.angle(Ep[:, e, a])), ws]
            res = []
            for n, w in enumerate(sets):
                w = w.astype(np.complex64).astype(np.complex128)
                bt = np.tensordot(w, Et.astype(np.complex128), axes=(0, 0))
                bp = np.tensordot(w, Ep.astype(np.complex128), axes=(0, 0))
                tsa = 4 * np.pi if n == 0 else total_solid_angle
                dt, dp, dtot = directivity_transform(bt, bp, az_range, elev_range, tsa)
                res.append((dt[e, a], dp[e, a], dtot[e, a]))
            col = {"Total": 2, "Etheta": 0, "Ephi": 1}[forming]
            best = int(np.argmax([r[col] for r in res]))
            out[e, a] = res[best]
    return out
