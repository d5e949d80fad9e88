"""Pins the oracle's ray-triangle predicates and the kernel1Dv3 loop.

Each check is independent of the oracle's own formula: closed-form geometry,
exact rational arithmetic (fractions) on random tiny scenes, and the reference's
documented window / culling rules.
"""
import numpy as np
import pytest

TRI_Z1 = np.array([0, 0, 1, 1, 0, 1, 0, 1, 1], np.float32)  # plane z = 1, normal e1 x e2 = +z


def _ray(oracle, o, t):
    d, L = oracle.charge(np.array([o], np.float32), np.array([t], np.float32))
    return np.array(o, np.float32), d[0], float(L[0])


def test_axis_aligned_hit_both_directions(oracle):
    # upward ray: A = e1 . (d x e2) = -1 -> back face: numba hits (two-sided), native culls
    o, d, L = _ray(oracle, [0.2, 0.2, 0.0], [0.2, 0.2, 2.0])
    h, t = oracle.hit_numba(o, d, L, TRI_Z1)
    assert h and t == pytest.approx(1.0, abs=1e-7)
    assert not oracle.hit_native(o, d, L, TRI_Z1)
    # downward ray: front face for both
    o, d, L = _ray(oracle, [0.2, 0.2, 2.0], [0.2, 0.2, 0.0])
    h, t = oracle.hit_numba(o, d, L, TRI_Z1)
    assert h and t == pytest.approx(1.0, abs=1e-7)
    assert oracle.hit_native(o, d, L, TRI_Z1)


@pytest.mark.parametrize("xy", [(-0.1, 0.3), (0.3, -0.1), (0.6, 0.6), (1.2, 0.0)])
def test_barycentric_misses(oracle, xy):
    o, d, L = _ray(oracle, [xy[0], xy[1], 2.0], [xy[0], xy[1], 0.0])
    assert not oracle.hit_numba(o, d, L, TRI_Z1)[0]
    assert not oracle.hit_native(o, d, L, TRI_Z1)


def test_parallel_ray_misses(oracle):
    o, d, L = _ray(oracle, [-1.0, 0.2, 1.0], [2.0, 0.2, 1.0])
    assert not oracle.hit_numba(o, d, L, TRI_Z1)[0]
    assert not oracle.hit_native(o, d, L, TRI_Z1)


def test_accept_window_edges(oracle):
    tri = np.array([-1, -1, 0, 3, -1, 0, -1, 3, 0], np.float32)  # plane z = 0, normal +z
    # numba window (2e-7, L - 2e-7) on the float64 t (both faces)
    for dz, expect in [(1.5e-7, False), (3e-7, True)]:
        o, d, L = _ray(oracle, [0.1, 0.1, dz], [0.1, 0.1, -0.5])   # t = dz
        assert oracle.hit_numba(o, d, L, tri)[0] is expect
        o, d, L = _ray(oracle, [0.1, 0.1, -dz], [0.1, 0.1, 0.5])   # back face, t = dz
        assert oracle.hit_numba(o, d, L, tri)[0] is expect
        o, d, L = _ray(oracle, [0.1, 0.1, 0.5], [0.1, 0.1, -dz])   # L - t = dz
        assert oracle.hit_numba(o, d, L, tri)[0] is expect
    # native window (2e-6, L - 2e-6), front faces only
    for dz, expect in [(1.5e-6, False), (3e-6, True)]:
        o, d, L = _ray(oracle, [0.1, 0.1, dz], [0.1, 0.1, -0.5])
        assert oracle.hit_native(o, d, L, tri) is expect
        o, d, L = _ray(oracle, [0.1, 0.1, 0.5], [0.1, 0.1, -dz])
        assert oracle.hit_native(o, d, L, tri) is expect
        o, d, L = _ray(oracle, [0.1, 0.1, -dz], [0.1, 0.1, 0.5])   # back face: culled
        assert not oracle.hit_native(o, d, L, tri)


def test_near_degenerate_determinant(oracle):
    # a sliver triangle: |A| below 1e-7 is a miss for numba, below 1e-6 for native
    tri = np.array([0, 0, 1, 1e-4, 0, 1, 0, 1e-4, 1], np.float32)  # |A| = 1e-8
    o, d, L = _ray(oracle, [2e-5, 2e-5, 2.0], [2e-5, 2e-5, 0.0])
    assert not oracle.hit_numba(o, d, L, tri)[0]
    tri = np.array([0, 0, 1, 1e-3, 0, 1, 0, 1e-3, 1], np.float32)  # |A| = 1e-6 (float32 ~)
    o, d, L = _ray(oracle, [2e-4, 2e-4, 2.0], [2e-4, 2e-4, 0.0])
    assert oracle.hit_numba(o, d, L, tri)[0]


def test_exact_rational_geometry_random(oracle):
    """Away from edges and window ends, the float32 predicates must agree with
    exact rational segment/triangle intersection."""
    rng = np.random.default_rng(11)
    decided = 0
    for _ in range(600):
        o = rng.uniform(-1, 1, 3).astype(np.float32)
        t = rng.uniform(-1, 1, 3).astype(np.float32)
        tri = rng.uniform(-1, 1, 9).astype(np.float32)
        res = oracle.exact_segment_triangle(o, t, tri)
        if res is None:
            continue
        margin, s = res
        o32, d, L = _ray(oracle, o, t)
        margin = float(margin)
        s = float(s)
        inside = margin > 1e-4 and 1e-4 < s < 1 - 1e-4
        outside = margin < -1e-4 or s < -1e-4 or s > 1 + 1e-4
        if not (inside or outside):
            continue
        decided += 1
        h, tt = oracle.hit_numba(o32, d, L, tri)
        assert h is inside
        if inside:
            assert tt == pytest.approx(s * L, rel=1e-5, abs=1e-6)
            e1 = tri[3:6].astype(np.float64) - tri[0:3]
            e2 = tri[6:9].astype(np.float64) - tri[0:3]
            front = np.dot(np.cross(e1, e2), t.astype(np.float64) - o) < 0
            assert oracle.hit_native(o32, d, L, tri) is bool(front)
        else:
            assert not oracle.hit_native(o32, d, L, tri)
    assert decided > 400


def test_kernel1dv3_closest_hit_and_order(oracle):
    near = np.array([-1, -1, 1, 3, -1, 1, -1, 3, 1], np.float32)
    far = np.array([-1, -1, 2, 3, -1, 2, -1, 3, 2], np.float32)
    o, d, L = _ray(oracle, [0.1, 0.1, 0.0], [0.1, 0.1, 3.0])
    ray = np.concatenate([o, d, [L]]).astype(np.float32).reshape(1, 7)
    for tris in (np.stack([near, far]), np.stack([far, near])):
        inter, dist = oracle.raycast_rays(ray, tris)
        assert inter[0] and dist[0] == pytest.approx(1.0, abs=1e-6)
    inter, dist = oracle.raycast_rays(ray, np.empty((0, 9), np.float32))
    assert not inter[0] and dist[0] == np.float32(L)


def test_dense_visibility_matches_single_predicate(oracle):
    rng = np.random.default_rng(5)
    src = rng.uniform(-1, 1, (7, 3)).astype(np.float32)
    snk = rng.uniform(-1, 1, (9, 3)).astype(np.float32) + np.float32(2.5)
    ctr = rng.uniform(0.0, 2.0, (40, 1, 3))
    tris = (ctr + rng.uniform(-0.4, 0.4, (40, 3, 3))).reshape(40, 9).astype(np.float32)
    hit, dist = oracle.visibility_dense(src, snk, tris)
    anyh, _ = oracle.visibility_dense(src, snk, tris, any_hit_only=True, want_dist=False)
    assert (hit == anyh).all()
    for s in range(7):
        for k in range(9):
            o, d, L = _ray(oracle, src[s], snk[k])
            ts = [oracle.hit_numba(o, d, L, tr) for tr in tris]
            hits = [t for h, t in ts if h]
            assert hit[s, k] == bool(hits)
            if hits:
                assert dist[s, k] == pytest.approx(min(hits), rel=1e-6)
            else:
                assert dist[s, k] == np.float32(L)
    assert 0 < hit.sum() < hit.size


def test_charge_unit_directions(oracle):
    rng = np.random.default_rng(3)
    o = rng.uniform(-50, 50, (200, 3)).astype(np.float32)
    t = rng.uniform(-50, 50, (200, 3)).astype(np.float32)
    d, L = oracle.charge(o, t)
    ref = np.linalg.norm(t.astype(np.float64) - o, axis=1)
    assert np.allclose(L, ref, rtol=3e-7)
    assert np.allclose(np.linalg.norm(d.astype(np.float64), axis=1), 1.0, atol=5e-7)
    assert np.allclose(d * L[:, None], t.astype(np.float64) - o, atol=2e-5)
