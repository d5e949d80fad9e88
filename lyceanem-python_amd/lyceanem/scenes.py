"""Seeded synthetic scenes of the benchmark configurations.

Each generator returns plain numpy arrays (float32 points, triangle tables as
(T, 9) float32, complex64 weights) so the same scene feeds the drop-in API, the
C-ABI and the CPU oracle.  Sizes can be scaled down for tests.

planar_array_c1    -- 32x32 lambda/2 array at 3 GHz, 2,000 random triangles,
                      91 x 181 far-field sinks at 1 km.
terrain_c2         -- 1000 x 500 quad (1,000,000 triangle) height field with
                      Gaussian-correlated heights, 1024 x 1024 sources at +30 m,
                      16,384 sinks at +1.5 m.
"""
from __future__ import annotations

import numpy as np

C0 = 299792458.0


def sphere_sinks(n_el=91, n_az=181, radius=1000.0, outward=True):
    """Far-field sinks: el-major, az-fastest grid (degrees) at `radius`."""
    el = np.radians(np.linspace(-90.0, 90.0, n_el))
    az = np.radians(np.linspace(-180.0, 180.0, n_az))
    azg, elg = np.meshgrid(az, el)
    p = np.stack([np.cos(elg) * np.cos(azg), np.cos(elg) * np.sin(azg), np.sin(elg)], -1)
    p = p.reshape(-1, 3)
    n = p if outward else -p
    return (radius * p).astype(np.float32), n.astype(np.float32)


def planar_array_c1(n=32, wavelength=0.1, n_tris=2000, n_el=91, n_az=181, radius=1000.0,
                    seed=1):
    """Config 1: n x n array at lambda/2 pitch centred at the origin in z=0,
    normals +z, x-polarised unit weights (Area 1); n_tris random triangles with
    vertices uniform in [-1,1]^2 x [0.05,1.05] m; far-field sinks."""
    pitch = wavelength / 2
    c = (np.arange(n) - (n - 1) / 2) * pitch
    xx, yy = np.meshgrid(c, c)
    src = np.stack([xx.ravel(), yy.ravel(), np.zeros(n * n)], 1).astype(np.float32)
    src_n = np.tile(np.array([[0, 0, 1]], np.float32), (n * n, 1))
    w = np.zeros((n * n, 3), np.complex64)
    w[:, 0] = 1.0
    rng = np.random.default_rng(seed)
    v = rng.uniform(0, 1, size=(n_tris, 3, 3))
    v[..., 0] = v[..., 0] * 2 - 1
    v[..., 1] = v[..., 1] * 2 - 1
    v[..., 2] = v[..., 2] + 0.05
    tris = v.reshape(n_tris, 9).astype(np.float32)
    snk, snk_n = sphere_sinks(n_el, n_az, radius)
    return dict(src=src, src_n=src_n, w=w, tris=tris, snk=snk, snk_n=snk_n,
                wavelength=wavelength, alpha=0.0, beta=2 * np.pi / wavelength)


def terrain_heights(nx=1001, ny=501, dx=1.0, sigma_h=5.0, corr=50.0, seed=2):
    """Vertex heights (ny, nx) of a Gaussian-correlated height field: white
    noise filtered with a Gaussian of std `corr` metres, rescaled to std sigma_h."""
    from scipy.ndimage import gaussian_filter

    rng = np.random.default_rng(seed)
    h = gaussian_filter(rng.standard_normal((ny, nx)), sigma=corr / dx, mode="wrap")
    h = (h - h.mean()) / h.std() * sigma_h
    return h


def terrain_surface_height(h, x, y, dx=1.0):
    """Height of the triangulated surface at (x, y): each quad is split along its
    (i,j)-(i+1,j+1) diagonal, as terrain_c2 triangulates it."""
    ny, nx = h.shape
    fx = np.clip(x / dx, 0, nx - 1 - 1e-9)
    fy = np.clip(y / dx, 0, ny - 1 - 1e-9)
    i = np.floor(fx).astype(np.int64)
    j = np.floor(fy).astype(np.int64)
    u = fx - i
    v = fy - j
    h00 = h[j, i]
    h10 = h[j, i + 1]
    h01 = h[j + 1, i]
    h11 = h[j + 1, i + 1]
    lower = u >= v  # triangle (00, 10, 11)
    z_lo = h00 + u * (h10 - h00) + v * (h11 - h10)
    z_up = h00 + v * (h01 - h00) + u * (h11 - h01)
    return np.where(lower, z_lo, z_up)


def terrain_c2(nqx=1000, nqy=500, size_x=1000.0, size_y=500.0, n_src_side=1024, n_snk=16384,
               src_alt=30.0, snk_alt=1.5, seed_h=2, seed_snk=3):
    """Config 2: height-field terrain of nqx x nqy quads (2 triangles each) over
    size_x x size_y metres; sources on an n_src_side^2 grid at src_alt above the
    surface, sinks uniform over the domain at snk_alt above the surface."""
    dx = size_x / nqx
    assert abs(size_y / nqy - dx) < 1e-9, "square quads expected"
    h = terrain_heights(nqx + 1, nqy + 1, dx=dx, seed=seed_h) if seed_h is not None else \
        np.zeros((nqy + 1, nqx + 1))
    ny, nx = h.shape
    X, Y = np.meshgrid(np.arange(nx) * dx, np.arange(ny) * dx)
    V = np.stack([X, Y, h], -1)  # (ny, nx, 3)
    p00 = V[:-1, :-1].reshape(-1, 3)
    p10 = V[:-1, 1:].reshape(-1, 3)
    p01 = V[1:, :-1].reshape(-1, 3)
    p11 = V[1:, 1:].reshape(-1, 3)
    t_lo = np.concatenate([p00, p10, p11], 1)
    t_up = np.concatenate([p00, p11, p01], 1)
    tris = np.stack([t_lo, t_up], 1).reshape(-1, 9).astype(np.float32)
    gx = (np.arange(n_src_side) + 0.5) * (size_x / n_src_side)
    gy = (np.arange(n_src_side) + 0.5) * (size_y / n_src_side)
    sx, sy = np.meshgrid(gx, gy)
    sx, sy = sx.ravel(), sy.ravel()
    src = np.stack([sx, sy, terrain_surface_height(h, sx, sy, dx) + src_alt], 1)
    rng = np.random.default_rng(seed_snk)
    kx = rng.uniform(0, size_x, n_snk)
    ky = rng.uniform(0, size_y, n_snk)
    snk = np.stack([kx, ky, terrain_surface_height(h, kx, ky, dx) + snk_alt], 1)
    return dict(tris=tris, src=src.astype(np.float32), snk=snk.astype(np.float32), heights=h,
                dx=dx)
