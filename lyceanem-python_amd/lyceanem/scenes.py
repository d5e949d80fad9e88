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
               src_alt=30.0, snk_alt=1.5, seed_h=2, seed_snk=3, sigma_h=5.0, corr=50.0):
    """Config 2: height-field terrain of nqx x nqy quads (2 triangles each) over
    size_x x size_y metres; sources on an n_src_side^2 grid at src_alt above the
    surface, sinks uniform over the domain at snk_alt above the surface."""
    dx = size_x / nqx
    assert abs(size_y / nqy - dx) < 1e-9, "square quads expected"
    h = terrain_heights(nqx + 1, nqy + 1, dx=dx, sigma_h=sigma_h, corr=corr, seed=seed_h) \
        if seed_h is not None else \
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


def clutter_c3(n_src_side=256, n_tris=50000, n_scat=200000, n_snk_side=64, freq=5.8e9,
               seed_w=4, seed_mesh=5, seed_scat=6, reflection=0.5, offset=1e-4):
    """Config 3: one-bounce scattering network at `freq`.

    Sources: n x n array at lambda/2 in z=0 centred at the origin, normals +z,
    x-polarised weights with seeded uniform random phase.  Mesh: n_tris random
    triangles (vertices = centre + U[-1.5,1.5]^3 m) with centres uniform in
    [-90,90] x [-90,90] x [20,200] m.  Scatterers: n_scat area-weighted uniform
    samples on the triangles, carrying the unit face normal and lifted `offset`
    metres along it, weight `reflection`.  Sinks: n x n grid in z=0 over
    x in [20,200], y in [-90,90] m, normals pointing away from the array."""
    lam = C0 / freq
    pitch = lam / 2
    c = (np.arange(n_src_side) - (n_src_side - 1) / 2) * pitch
    xx, yy = np.meshgrid(c, c)
    S = n_src_side * n_src_side
    src = np.stack([xx.ravel(), yy.ravel(), np.zeros(S)], 1).astype(np.float32)
    src_n = np.tile(np.array([[0, 0, 1]], np.float32), (S, 1))
    rng = np.random.default_rng(seed_w)
    w = np.zeros((S, 3), np.complex64)
    w[:, 0] = np.exp(1j * rng.uniform(0, 2 * np.pi, S))
    rng = np.random.default_rng(seed_mesh)
    ctr = np.stack([rng.uniform(-90, 90, n_tris), rng.uniform(-90, 90, n_tris),
                    rng.uniform(20, 200, n_tris)], 1)
    v = ctr[:, None, :] + rng.uniform(-1.5, 1.5, (n_tris, 3, 3))
    tris = v.reshape(-1, 9).astype(np.float32)
    t64 = tris.astype(np.float64).reshape(-1, 3, 3)
    cr = np.cross(t64[:, 1] - t64[:, 0], t64[:, 2] - t64[:, 0])
    area = 0.5 * np.linalg.norm(cr, axis=1)
    fn = cr / np.linalg.norm(cr, axis=1, keepdims=True)
    rng = np.random.default_rng(seed_scat)
    pick = rng.choice(n_tris, n_scat, p=area / area.sum())
    r1 = np.sqrt(rng.uniform(0, 1, n_scat))
    r2 = rng.uniform(0, 1, n_scat)
    a, b, cc = t64[pick, 0], t64[pick, 1], t64[pick, 2]
    pts = (1 - r1)[:, None] * a + (r1 * (1 - r2))[:, None] * b + (r1 * r2)[:, None] * cc
    scat_n = fn[pick]
    scat = (pts + offset * scat_n).astype(np.float32)
    g = np.linspace(0, 1, n_snk_side)
    kx, ky = np.meshgrid(20 + 180 * g, -90 + 180 * g)
    K = n_snk_side * n_snk_side
    snk = np.stack([kx.ravel(), ky.ravel(), np.zeros(K)], 1).astype(np.float32)
    snk_n = (snk / np.linalg.norm(snk, axis=1, keepdims=True)).astype(np.float32)
    return dict(src=src, src_n=src_n, w=w, tris=tris, scat=scat,
                scat_n=scat_n.astype(np.float32),
                scat_w=np.full((n_scat, 3), reflection, np.complex64), snk=snk, snk_n=snk_n,
                wavelength=lam, alpha=0.0, beta=2 * np.pi / lam)


def conformal_c4(n_az_el=64, n_z_el=64, radius=0.5, wavelength=0.03, n_el=181, n_az=361,
                 far=1000.0, seed_w=7, standoff=1e-3, n_side=128, n_ring=64):
    """Config 4: conformal array on a cylinder for beam synthesis.

    Elements: n_az_el around the full circumference x n_z_el along z at lambda/2
    pitch, `standoff` outside a cylinder of `radius`, radial normals,
    z-polarised unit weights with seeded random phase.  Blocking mesh: the
    cylinder side (n_side facets x n_ring rings, 2 triangles each, plus end caps)
    and a radial fin at azimuth 90 deg (1 m out, 0.8 m tall); ~20k triangles.
    Sinks: n_el x n_az (el-major, az-fastest) far-field grid at `far` metres."""
    phi = 2 * np.pi * np.arange(n_az_el) / n_az_el
    z = (np.arange(n_z_el) - (n_z_el - 1) / 2) * (wavelength / 2)
    pg, zg = np.meshgrid(phi, z)
    pg, zg = pg.ravel(), zg.ravel()
    n = np.stack([np.cos(pg), np.sin(pg), np.zeros_like(pg)], 1)
    pos = (n * (radius + standoff) + np.stack([0 * pg, 0 * pg, zg], 1)).astype(np.float32)
    rng = np.random.default_rng(seed_w)
    S = len(pg)
    w = np.zeros((S, 3), np.complex64)
    w[:, 2] = np.exp(1j * rng.uniform(0, 2 * np.pi, S))
    h = n_z_el * wavelength / 2 * 1.25
    a = 2 * np.pi * np.arange(n_side + 1) / n_side
    zz = np.linspace(-h / 2, h / 2, n_ring + 1)
    P = np.stack([radius * np.cos(a)[None, :].repeat(n_ring + 1, 0),
                  radius * np.sin(a)[None, :].repeat(n_ring + 1, 0),
                  zz[:, None].repeat(n_side + 1, 1)], -1)
    q00, q10, q01, q11 = P[:-1, :-1], P[:-1, 1:], P[1:, :-1], P[1:, 1:]
    side = np.concatenate([np.concatenate([q00, q10, q11], -1).reshape(-1, 9),
                           np.concatenate([q00, q11, q01], -1).reshape(-1, 9)])
    caps = []
    for zc in (-h / 2, h / 2):
        c = np.array([0.0, 0.0, zc])
        for i in range(n_side):
            p0 = np.array([radius * np.cos(a[i]), radius * np.sin(a[i]), zc])
            p1 = np.array([radius * np.cos(a[i + 1]), radius * np.sin(a[i + 1]), zc])
            caps.append(np.concatenate([c, p0, p1]))
    fr = np.linspace(radius, radius + 1.0, 61)
    fz = np.linspace(-0.4, 0.4, 29)
    F = np.stack([0 * fr[None, :].repeat(29, 0), fr[None, :].repeat(29, 0),
                  fz[:, None].repeat(61, 1)], -1)
    f00, f10, f01, f11 = F[:-1, :-1], F[:-1, 1:], F[1:, :-1], F[1:, 1:]
    fin = np.concatenate([np.concatenate([f00, f10, f11], -1).reshape(-1, 9),
                          np.concatenate([f00, f11, f01], -1).reshape(-1, 9)])
    tris = np.concatenate([side, np.asarray(caps), fin]).astype(np.float32)
    snk, snk_n = sphere_sinks(n_el, n_az, far)
    el = np.repeat(np.linspace(-90.0, 90.0, n_el), n_az)
    az = np.tile(np.linspace(-180.0, 180.0, n_az), n_el)
    in_beam = (np.abs(az - 30.0) <= 5.0) & (np.abs(el) <= 5.0)
    return dict(src=pos, src_n=n.astype(np.float32), w=w, tris=tris, snk=snk, snk_n=snk_n,
                in_beam=in_beam, wavelength=wavelength, alpha=0.0, beta=2 * np.pi / wavelength)


def batched_c5(n_src_side=128, n_el=128, n_az=128, n_freq=32, f_lo=2e9, f_hi=4e9, n_sets=64,
               far=1000.0, seed=10):
    """Config 5: batched excitations for the G . W contraction.

    Sources: n x n grid at lambda_min/2 (f_hi) in z=0, z-polarised.  Sinks:
    n_el x n_az far-field grid (el-major, az-fastest) at `far` metres.  n_freq
    wavenumbers evenly spread over [f_lo, f_hi].  W[f, s, b] = exp(-j k_f p_s . u_b)
    steers set b towards one of n_sets seeded random upper-hemisphere directions."""
    lam_min = C0 / f_hi
    pitch = lam_min / 2
    c = (np.arange(n_src_side) - (n_src_side - 1) / 2) * pitch
    xx, yy = np.meshgrid(c, c)
    S = n_src_side * n_src_side
    src = np.stack([xx.ravel(), yy.ravel(), np.zeros(S)], 1).astype(np.float32)
    snk, snk_n = sphere_sinks(n_el, n_az, far)
    freqs = np.linspace(f_lo, f_hi, n_freq)
    kf = 2 * np.pi * freqs / C0
    rng = np.random.default_rng(seed)
    th = np.arccos(rng.uniform(0.0, 1.0, n_sets))
    ph = rng.uniform(-np.pi, np.pi, n_sets)
    u = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    proj = src.astype(np.float64) @ u.T  # (S, B)
    W = np.exp(-1j * kf[:, None, None] * proj[None]).astype(np.complex64)
    return dict(src=src, snk=snk, snk_n=snk_n, pol=np.array([0, 0, 1], np.float32),
                wavenumbers=kf, W=W, directions=u)


def time_domain_td(n_src_side=64, n_snk=4096, freq=3e9, fs=40e9, n_samples=65536, n_exc=256,
                   r_min=50.0, r_max=500.0, seed=13):
    """Time-domain workload: an n x n array at lambda/2 in z=0 (normals +z,
    x-polarised, weight = Area = (lambda/2)^2), n_snk sinks 50-500 m away in the
    upper half-space (normals back towards the array), every source -> sink
    line-of-sight path, a Gaussian-windowed sine pulse at `freq` sampled at fs
    (n_exc samples) and an n_samples window."""
    lam = C0 / freq
    pitch = lam / 2
    c = (np.arange(n_src_side) - (n_src_side - 1) / 2) * pitch
    xx, yy = np.meshgrid(c, c)
    S = n_src_side * n_src_side
    src = np.stack([xx.ravel(), yy.ravel(), np.zeros(S)], 1).astype(np.float32)
    rng = np.random.default_rng(seed)
    th = np.arccos(rng.uniform(0.05, 1.0, n_snk))
    ph = rng.uniform(-np.pi, np.pi, n_snk)
    r = rng.uniform(r_min, r_max, n_snk)
    u = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    snk = (u * r[:, None]).astype(np.float32)
    snk_n = (-u).astype(np.float32)
    w = np.zeros((S, 3), np.complex64)
    w[:, 0] = pitch * pitch
    pos = np.concatenate([src, snk])
    nrm = np.concatenate([np.tile(np.array([0, 0, 1], np.float32), (S, 1)), snk_n])
    weights = np.concatenate([w, np.ones((n_snk, 3), np.complex64)])
    s_idx, k_idx = np.meshgrid(np.arange(S), np.arange(n_snk), indexing="ij")
    index = np.stack([s_idx.ravel() + 1, k_idx.ravel() + S + 1], 1).astype(np.int32)
    t = (np.arange(n_exc) - n_exc / 2) / fs
    sig = n_exc / (8 * fs)
    exc = np.sin(2 * np.pi * freq * t) * np.exp(-0.5 * (t / sig) ** 2)
    return dict(pos=pos, nrm=nrm, w=weights, index=index, exc=exc, fs=fs, n_samples=n_samples,
                beta=2 * np.pi / lam, S=S, K=n_snk)


def beamforming_bf(n_side=32, freq=3e9, n_el=91, n_az=181, coupling=0.05, seed=14):
    """Directivity-map workload: an n x n array at lambda/2 in z = 0 with
    element patterns Etheta = max(cos(angle off +z), 0.05) e^{jk p.u} and
    Ephi = 0.5 cos(az) e^{jk p.u} plus seeded random coupling, on an
    n_el x n_az grid (el-major, az-fastest); every grid point is a command
    angle."""
    lam = C0 / freq
    g = (np.arange(n_side) - (n_side - 1) / 2) * lam / 2
    X, Y = np.meshgrid(g, g)
    pts = np.stack([X.ravel(), Y.ravel(), np.zeros(n_side * n_side)], 1)
    az = np.linspace(-180.0, 180.0, n_az)
    el = np.linspace(-90.0, 90.0, n_el)
    A, E = np.meshgrid(np.radians(az), np.radians(el))
    u = np.stack([np.cos(E) * np.cos(A), np.cos(E) * np.sin(A), np.sin(E)], -1).reshape(-1, 3)
    phase = np.exp(1j * (2 * np.pi / lam) * (pts @ u.T)).astype(np.complex64)
    rng = np.random.default_rng(seed)
    n, P = len(pts), n_el * n_az
    ef = np.clip(u[:, 2], 0.05, None).astype(np.float32)
    noise = lambda: (coupling * (rng.standard_normal((n, P), np.float32)
                                 + 1j * rng.standard_normal((n, P), np.float32))).astype(np.complex64)
    Et = (ef[None, :] * phase + noise()).reshape(n, n_el, n_az)
    Ep = ((0.5 * np.cos(A).ravel()).astype(np.float32)[None, :] * phase + noise()).reshape(
        n, n_el, n_az)
    return dict(points=pts, Et=Et, Ep=Ep, az=az, el=el, wavelength=lam)
