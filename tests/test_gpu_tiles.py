"""GPU tests of the native-branch drop-ins: lyceanem.em (tile binning, per-pair
tile / brute-force networks) and models.acceleration_structures."""
import numpy as np
import pytest

from lyceanem.mesh import Mesh

pytestmark = pytest.mark.gpu


def _terrain_mesh(n=12, seed=2):
    """Height field with x vertical (the orientation the tile grid bins in y, z)."""
    rng = np.random.default_rng(seed)
    y, z = np.meshgrid(np.linspace(-1, 1, n + 1), np.linspace(-1, 1, n + 1))
    x = 0.15 * rng.normal(size=y.shape)
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    idx = np.arange((n + 1) ** 2).reshape(n + 1, n + 1)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, :-1].ravel(), idx[1:, 1:].ravel()
    tri = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)])
    tri = tri[rng.random(len(tri)) < 0.3]  # holes, so part of the rays pass
    return Mesh(points=pts, cells=[("triangle", tri)])


@pytest.mark.parametrize("n", [12, 90])
def test_tile_structure_network_vs_oracle(oracle, n):
    from lyceanem.models.acceleration_structures import Tile_acceleration_structure

    m = _terrain_mesh(n=n)
    rng = np.random.default_rng(4)
    S, K = 40, 70
    src = np.stack([rng.uniform(0.5, 1.5, S), rng.uniform(-0.8, 0.8, S), rng.uniform(-0.8, 0.8, S)], 1)
    snk = np.stack([rng.uniform(-1.5, -0.5, K), rng.uniform(-0.8, 0.8, K), rng.uniform(-0.8, 0.8, K)], 1)
    w = (rng.normal(size=(S, 3)) + 1j * rng.normal(size=(S, 3))).astype(np.complex64)
    src_m = Mesh(points=src, point_data={"Normals": np.tile([-1.0, 0, 0], (S, 1)),
                                         "ex": w[:, 0], "ey": w[:, 1], "ez": w[:, 2]})
    snk_m = Mesh(points=snk, point_data={"Normals": np.tile([1.0, 0, 0], (K, 1))})
    t = Tile_acceleration_structure(m, 1)
    beta = 2 * np.pi / 0.2
    net = t.calculate_scattering(src_m, snk_m, 0.0, beta, 0.2, self_to_self=False, chunk_count=3)
    assert net.shape == (S, K, 3) and net.dtype == np.complex64
    tris9 = np.asarray(m.points, np.float32)[m.cells[0].data].reshape(-1, 9)
    # self_to_self=False skips chunk-local source row == sink row: chunks of
    # floor(40 / 3) = 13 rows, the last running to 40
    hit = np.concatenate([
        oracle.visibility_dense(src[a:b], snk, tris9, semantics=1, any_hit_only=True,
                                skip_self=True, want_dist=False)[0]
        for a, b in ((0, 13), (13, 26), (26, 40))])
    ref = oracle.field_los(src, w, snk, snk_m.point_data["Normals"], 0.0, beta, hit=hit,
                           elements=True)
    assert 0.05 < hit.mean() < 0.95
    assert np.linalg.norm(net - ref) / np.linalg.norm(ref) < 1e-4
    assert (net[hit] == 0).all()


def test_brute_force_structure_returns_network(oracle):
    from lyceanem.models.acceleration_structures import Brute_Force_acceleration_structure

    m = _terrain_mesh(n=6)
    rng = np.random.default_rng(5)
    S, K = 9, 11
    src = np.stack([np.full(S, 1.0), rng.uniform(-0.8, 0.8, S), rng.uniform(-0.8, 0.8, S)], 1)
    snk = np.stack([np.full(K, -1.0), rng.uniform(-0.8, 0.8, K), rng.uniform(-0.8, 0.8, K)], 1)
    w = np.ones((S, 3), np.complex64)
    src_m = Mesh(points=src, point_data={"Normals": np.tile([-1.0, 0, 0], (S, 1)),
                                         "ex": w[:, 0], "ey": w[:, 1], "ez": w[:, 2]})
    snk_m = Mesh(points=snk, point_data={"Normals": np.tile([1.0, 0, 0], (K, 1))})
    net = Brute_Force_acceleration_structure(m).calculate_scattering(
        src_m, snk_m, 0.0, 2 * np.pi, 1.0, self_to_self=True)
    tris9 = np.asarray(m.points, np.float32)[m.cells[0].data].reshape(-1, 9)
    hit, _ = oracle.visibility_dense(src, snk, tris9, semantics=1, any_hit_only=True,
                                     want_dist=False)
    ref = oracle.field_los(src, w, snk, snk_m.point_data["Normals"], 0.0, 2 * np.pi, hit=hit,
                           elements=True)
    assert np.abs(ref).max() > 0
    assert np.linalg.norm(net - ref) / np.linalg.norm(ref) < 1e-4


def test_em_module_validates_like_the_reference():
    from lyceanem import em

    with pytest.raises(RuntimeError, match="3 columns"):
        em.calculate_scattering_brute_force(np.zeros((3, 2)), np.zeros((2, 3)),
                                            np.zeros((1, 3), int), np.zeros((3, 3)), 1.0,
                                            *[np.zeros(3)] * 6, np.zeros((5, 3)), 0.0, 1.0, True)
    with pytest.raises(RuntimeError, match="ex_imag"):
        em.calculate_scattering_brute_force(np.zeros((3, 3)), np.zeros((2, 3)),
                                            np.zeros((1, 3), int), np.zeros((3, 3)), 1.0,
                                            np.zeros(3), np.zeros(2), *[np.zeros(3)] * 4,
                                            np.zeros((5, 3)), 0.0, 1.0, True)
