"""Pins the oracle's restatement of the reference's tile acceleration
(Tile_acceleration_structure + the tile DDA of raycasting_accelerated.cuh),
which bench.py times as a second CPU baseline for the terrain launch: on
height-field terrains, where binning triangles by their vertices' tiles is
complete, its any-hit masks equal the brute force over every triangle; the
binning's known gap (a triangle whose vertices fall in two diagonal tiles is
not binned into the other two tiles it covers) is reproduced; segments above
the structure test nothing."""
import numpy as np
import pytest


@pytest.mark.parametrize("sigma,alt,corr,n_cells", [(8.0, 5.0, 20.0, 25), (10.0, 2.0, 15.0, 50)])
def test_tiles_equal_brute_force_on_terrain(oracle, sigma, alt, corr, n_cells):
    from lyceanem import scenes

    sc = scenes.terrain_c2(nqx=200, nqy=100, size_x=200.0, size_y=100.0, n_src_side=64,
                           n_snk=512, sigma_h=sigma, src_alt=alt, corr=corr)
    tiles = oracle.Tiles(sc["tris"], up=2, n_cells=n_cells)
    assert tiles.refs >= len(sc["tris"])
    rng = np.random.default_rng(n_cells)
    n = 20000
    pairs = np.stack([rng.integers(0, len(sc["src"]), n), rng.integers(0, len(sc["snk"]), n)], 1)
    hb, _ = oracle.visibility_sampled(sc["src"], sc["snk"], pairs, sc["tris"], any_hit_only=True)
    assert 0.2 < hb.mean() < 0.9
    assert np.array_equal(tiles.cast(sc["src"], sc["snk"], pairs), hb)
    # rays the other way (sink -> source: the ascending branch of the tile DDA)
    hr, _ = oracle.visibility_sampled(sc["snk"], sc["src"], pairs[:, ::-1], sc["tris"],
                                      any_hit_only=True)
    assert np.array_equal(tiles.cast(sc["snk"], sc["src"], pairs[:, ::-1]), hr)


def test_tiles_diagonal_binning_gap(oracle):
    # anchors fix the bounds to [0, 4] x [0, 4] (tile 4.004 / 4 = 1.001 m)
    anchors = np.array([[0, 0, 0, 0.01, 0, 0, 0, 0.01, 0],
                        [4, 4, 0, 3.99, 4, 0, 4, 3.99, 0]], np.float32)
    # vertices in tiles (0, 0), (1, 1), (1, 1): the triangle also covers parts
    # of tiles (1, 0) and (0, 1), where the reference does not bin it
    tri = np.array([[0.5, 0.5, 0, 1.9, 1.2, 0, 1.2, 1.9, 0]], np.float32)
    tiles = oracle.Tiles(np.concatenate([anchors, tri]), up=2, n_cells=4)
    assert (tiles.na, tiles.nb) == (4, 4)
    tris = np.concatenate([anchors, tri])
    # vertical rays through (1.1, 0.85) -- in tile (1, 0) -- and (1.5, 1.5)
    src = np.array([[1.1, 0.85, 5.0], [1.5, 1.5, 5.0]], np.float32)
    snk = np.array([[1.1, 0.85, -5.0], [1.5, 1.5, -5.0]], np.float32)
    pairs = np.array([[0, 0], [1, 1]])
    hb, _ = oracle.visibility_sampled(src, snk, pairs, tris, any_hit_only=True)
    assert hb.tolist() == [True, True]
    assert tiles.cast(src, snk, pairs).tolist() == [False, True]


def test_tiles_segment_above_structure_tests_nothing(oracle):
    # one tile (the reference needs triangles no wider than a tile)
    tri = np.array([[0, 0, 0, 10, 0, 0, 0, 10, 0], [10, 10, 0, 0, 10, 0, 10, 0, 0]], np.float32)
    tiles = oracle.Tiles(tri, up=2, n_cells=1)
    src = np.array([[1, 1, 3.0], [1, 1, 3.0]], np.float32)
    snk = np.array([[9, 9, 2.0], [9, 9, -2.0]], np.float32)
    pairs = np.array([[0, 0], [1, 1]])
    assert tiles.cast(src, snk, pairs).tolist() == [False, True]
    assert tiles.cast(src, snk, pairs, sem=1).tolist() == [False, True]
