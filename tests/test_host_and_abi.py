"""CPU tests: reference golden fixtures, host index logic of the drop-in package,
and the C-ABI library's exports (no compute calls: there is no GPU here)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")


# --------------------------------------------------------------------------
# reference fixtures
# --------------------------------------------------------------------------

def test_convert_triangles_golden():
    """tests/test_ray_functions.py:9-24 of the reference: convertTriangles of
    receive_horn.ply equals triangle_type_array_ref.npy."""
    from lyceanem import mesh
    from lyceanem.raycasting import rayfunctions as RF
    from lyceanem.base_types import triangle_t

    ref = np.load(os.path.join(GOLDEN, "triangle_type_array_ref.npy"), allow_pickle=False)
    m = mesh.read_ply(os.path.join(GOLDEN, "receive_horn.ply"))
    tri = RF.convertTriangles(m)
    assert tri.dtype == triangle_t and ref.dtype.names == triangle_t.names
    assert tri.shape == ref.shape == (12,)
    for name in triangle_t.names:
        assert np.array_equal(tri[name], ref[name]), name


@pytest.mark.parametrize("angle,e_in,e_out", [
    (0, [1, 0, 0], [0, 1, 0]), (0, [0, 1, 0], [0, 0, 1]), (0, [0, 0, 1], [1, 0, 0]),
    (-90, [1, 0, 0], [-1, 0, 0]), (-90, [0, 1, 0], [0, 0, 1]), (-90, [0, 0, 1], [0, 1, 0]),
])
def test_vector_mapping_known_answers(angle, e_in, e_out):
    """Known answers of tests/test_empropagation.py:8-95 (normal (1,0,0), rotation
    about z by 0 / -90 degrees, atol 1e-12)."""
    from scipy.spatial.transform import Rotation as R
    from lyceanem.electromagnetics.empropagation import vector_mapping

    rot = R.from_euler("z", angle, degrees=True).as_matrix()
    out = vector_mapping(np.array(e_in, np.complex64), np.array([1.0, 0, 0]), rot)
    np.testing.assert_allclose(out, np.array(e_out, np.complex64), atol=1e-12)


# --------------------------------------------------------------------------
# host index semantics: product vs oracle restatement
# --------------------------------------------------------------------------

def test_create_model_index_matches_oracle(oracle):
    from lyceanem.raycasting import rayfunctions as RF

    src = np.arange(1, 5).reshape(4, 1)
    snk = np.arange(5, 8).reshape(3, 1)
    sct = np.arange(8, 10).reshape(2, 1)
    for a, b, c in [(src, snk, sct), (src, snk, np.empty((0, 1))),
                    (np.array([[1, 8], [2, 9], [3, 8]]), snk, np.empty((0, 0)))]:
        got = RF.create_model_index(a, b, c)
        exp = oracle.create_model_index(a, b, c)
        assert got.dtype == np.int32 and np.array_equal(got, exp)
    # self rows are dropped (scatterer -> itself)
    got = RF.create_model_index(np.array([[1, 8], [1, 9]]), snk, sct)
    assert not (got[:, -2] == got[:, -1]).any()
    assert got.shape[0] == 2 * 5 - 2


def test_ray_hits_split(oracle):
    from lyceanem.raycasting import rayfunctions as RF

    idx = np.array([[1, 5], [1, 8], [2, 6], [2, 9], [3, 7]], np.int32)
    inter = np.array([False, False, True, False, False])
    sinks = np.arange(5, 8).reshape(3, 1)
    f, g = RF.rayHits1Dv2(inter, idx, sinks)
    f2, g2 = oracle.ray_hits(inter, idx, sinks)
    assert np.array_equal(f, f2) and np.array_equal(g, g2)
    # literal split (rayfunctions.py:1845-1878): the blocked row [2, 6] is
    # dropped, rows ending on a sink (5..7) are final, the rest continue
    assert np.array_equal(f2, [[1, 8], [2, 9]])
    assert np.array_equal(g2, [[1, 5], [3, 7]])


def test_targetting_index_and_slots(oracle):
    from lyceanem.electromagnetics import empropagation as EM

    fi = np.array([[1, 4, 0], [2, 5, 0], [1, 7, 4], [2, 8, 5], [3, 6, 0]], np.int32)
    d, sc = EM.targettingindex(fi)
    d2, sc2 = oracle.targetting_index(fi)
    assert np.array_equal(d, d2) and np.array_equal(sc, sc2)
    # literal answer of empropagation.py:3394-3429: line-of-sight rows first
    # (columns 0, 1), then the scattered rows (columns 0, 2); depth 2 marks the
    # rows with a scatterer, in input order
    assert np.array_equal(d2, [[1, 4], [2, 5], [3, 6], [1, 4], [2, 5]])
    assert np.array_equal(sc2, [1, 1, 2, 2, 1])
    d2, sc2 = oracle.targetting_index(fi[:, :2])
    assert np.array_equal(d2, fi[:, :2]) and np.array_equal(sc2, np.ones(5))
    s, k = EM.path_slots(fi, 3, 3)
    assert np.array_equal(s, [0, 1, 0, 1, 2])
    assert np.array_equal(k, [0, 1, 0, 1, 2])


def test_spherical_field_order_and_angles(oracle):
    from lyceanem.geometry.targets import spherical_field

    az = np.linspace(-180, 180, 7)
    el = np.linspace(-90, 90, 5)
    m = spherical_field(az, el, outward_normals=True, field_radius=2.0)
    p = m.points.reshape(5, 7, 3)
    # el-major, az-fastest; z depends on el only
    assert np.allclose(p[:, :, 2], 2 * np.sin(np.radians(el))[:, None])
    assert np.allclose(np.linalg.norm(m.points, axis=1), 2.0)
    assert np.allclose(m.point_data["Normals"], m.points / 2.0)
    th, ph, r = oracle.theta_phi(m.points)
    assert np.allclose(m.point_data["theta_(Radians)"], th)
    assert np.allclose(m.point_data["phi_(Radians)"], ph)
    # closed form of theta_phi_r (geometryfunctions.py:447-470) on this grid:
    # theta = pi/2 - el, phi = az (az = +-180 both map to atan2's +-pi; the
    # poles have no azimuth), r = the field radius
    elg, azg = np.meshgrid(np.radians(el), np.radians(az), indexing="ij")
    assert np.allclose(th.reshape(5, 7), np.pi / 2 - elg, atol=1e-6)
    assert np.allclose(r, 2.0)
    inner = (slice(1, 4), slice(1, 6))
    assert np.allclose(ph.reshape(5, 7)[inner], azg[inner], atol=1e-6)
    assert np.allclose(np.abs(ph.reshape(5, 7)[1:4, [0, 6]]), np.pi, atol=1e-6)
    inward = spherical_field(az, el, outward_normals=False)
    assert np.allclose(inward.point_data["Normals"], -inward.points)


def test_etheta_ephi_of_pure_polarisations(oracle):
    """Spherical-unit-vector identities of the calculate_farfield conversion."""
    th = np.radians(np.array([30.0, 60.0, 120.0]))
    ph = np.radians(np.array([10.0, 200.0, -45.0]))
    that = np.stack([np.cos(th) * np.cos(ph), np.cos(th) * np.sin(ph), -np.sin(th)], 1)
    phat = np.stack([-np.sin(ph), np.cos(ph), np.zeros(3)], 1)
    E = 2.0 * that + 3.0j * phat
    et, ep = oracle.etheta_ephi(E[:, 0], E[:, 1], E[:, 2], th, ph)
    assert np.allclose(et, 2.0) and np.allclose(ep, 3.0j)


def test_scene_generators_shapes():
    from lyceanem import scenes

    c1 = scenes.planar_array_c1(n=4, n_tris=10, n_el=5, n_az=7)
    assert c1["src"].shape == (16, 3) and c1["tris"].shape == (10, 9)
    assert c1["snk"].shape == (35, 3)
    c2 = scenes.terrain_c2(nqx=20, nqy=10, size_x=20.0, size_y=10.0, n_src_side=8, n_snk=16)
    assert c2["tris"].shape == (400, 9)
    # sinks sit snk_alt above the triangulated surface
    h = scenes.terrain_surface_height(c2["heights"], c2["snk"][:, 0], c2["snk"][:, 1])
    assert np.allclose(c2["snk"][:, 2] - h, 1.5, atol=1e-4)


# --------------------------------------------------------------------------
# C-ABI: the library loads and exports every declared entry point
# --------------------------------------------------------------------------

def _declared_symbols():
    text = open(os.path.join(ROOT, "include", "lem.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(lem_[a-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_symbol():
    from lyceanem import _native

    lib = _native.load_library()
    declared = _declared_symbols()
    assert "lem_visibility_dense" in declared and "lem_field_los" in declared
    for name in declared:
        assert hasattr(lib, name), name
    # the ctypes binding covers exactly the declared entry points
    assert sorted(_native.SIGNATURES) == declared
    assert lib.lem_abi_version() == _native.ABI_VERSION


def test_library_rejects_bad_arguments_without_gpu():
    """Argument validation happens before any HIP call."""
    from lyceanem import _native

    lib = _native.load_library()
    rc = lib.lem_visibility_dense(None, 4, None, 4, None, 0, 7, 0, None, None, None)
    assert rc == -1 and b"semantics" in lib.lem_last_error()
    rc = lib.lem_field_paths(None, None, None, None, 1, None, 5, None, None, None, 1, 0.0, 1.0,
                             None, None, None)
    assert rc == -1 and b"columns" in lib.lem_last_error()
    assert lib.lem_field_los_workspace(100, 1000, 0) > 0
    assert lib.lem_field_los_workspace(100, 1000, _native.FIELD_ELEMENTS) == 0
    # the allocation-free variant: its sink limit follows from the 16 MiB scratch,
    # and sizes beyond it are rejected before any HIP call
    kmax = lib.lem_field_los_noalloc_max_sinks()
    assert kmax == (16 << 20) // 4 // 12
    rc = lib.lem_field_los_noalloc(None, None, 10, None, None, None, kmax + 1, None, 0.0, 1.0, 0,
                                   None, None)
    assert rc == -1
    rc = lib.lem_field_los_noalloc(1, 1, 10, 1, 1, None, kmax + 1, None, 0.0, 1.0, 0, 1, None)
    assert rc == -1 and b"scratch" in lib.lem_last_error()
    # the one-pass loss: no sinks, missing pointers, a workspace below its size
    rc = lib.lem_db_mask_loss(None, None, 0, -25.0, 0.0, 1e-30, None, None, None, 0, None)
    assert rc == -1 and b"no sinks" in lib.lem_last_error()
    rc = lib.lem_db_mask_loss(None, None, 10, -25.0, 0.0, 1e-30, None, None, None, 0, None)
    assert rc == -1 and b"null pointer" in lib.lem_last_error()
    wsb = lib.lem_db_mask_loss_workspace(10)
    assert wsb >= 256 + 4
    rc = lib.lem_db_mask_loss(1, 1, 10, -25.0, 0.0, 1e-30, 1, None, 1, wsb - 1, None)
    assert rc != 0 and b"workspace" in lib.lem_last_error()


def test_product_fails_loudly_without_gpu():
    import torch
    from lyceanem import _native

    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    with pytest.raises(_native.LemError):
        _native.visibility_dense(np.zeros((1, 3)), np.ones((1, 3)), np.zeros((0, 9)))


# --------------------------------------------------------------------------
# tile binning of the native branch (host code)
# --------------------------------------------------------------------------

def _terrain_mesh_x_up(n=12, seed=2):
    from lyceanem.mesh import Mesh

    rng = np.random.default_rng(seed)
    y, z = np.meshgrid(np.linspace(-1, 1, n + 1), np.linspace(-1, 1, n + 1))
    x = 0.15 * rng.normal(size=y.shape)
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    idx = np.arange((n + 1) ** 2).reshape(n + 1, n + 1)
    a, b, c, d = (idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, :-1].ravel(),
                  idx[1:, 1:].ravel())
    tri = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)])
    return Mesh(points=pts, cells=[("triangle", tri)])


def test_tile_binning_covers_vertex_cells():
    """terrain_acceleration_build.cuh:57-205: each triangle sits in the tiles of
    its vertices, plus the fourth tile of an L-shaped triple."""
    from lyceanem.models.acceleration_structures import Tile_acceleration_structure

    m = _terrain_mesh_x_up()
    t = Tile_acceleration_structure(m, 4)
    tri = m.cells[0].data
    assert t.bin_counts.shape == (t.y_cells_count * t.z_cells_count,)
    assert t.binned_triangles_count == t.bin_counts.sum()
    bt = t.binned_triangles.reshape(-1, 3)
    assert {tuple(r) for r in bt} == {tuple(r) for r in tri}
    p = np.asarray(m.points, np.float32)
    cy = np.floor((p[:, 1] - np.float32(t.min_y)) / np.float32(t.tile_size)).astype(int)
    cz = np.floor((p[:, 2] - np.float32(t.min_z)) / np.float32(t.tile_size)).astype(int)
    starts = np.concatenate([[0], np.cumsum(t.bin_counts)])
    for i, tr in enumerate(tri[:60]):
        cells = {cy[v] * t.z_cells_count + cz[v] for v in tr}
        if len(cells) == 3:
            ys = {c // t.z_cells_count for c in cells}
            zs = {c % t.z_cells_count for c in cells}
            cells |= {y * t.z_cells_count + z for y in ys for z in zs}
        for c in cells:
            assert tuple(tr) in {tuple(r) for r in bt[starts[c]:starts[c + 1]]}
        assert sum(tuple(tr) in {tuple(r) for r in bt[starts[c]:starts[c + 1]]}
                   for c in range(len(t.bin_counts))) == len(cells)


@pytest.mark.parametrize("S,chunks", [(40, 3), (10, 4), (3, 5), (7, 1), (8, 2), (0, 2)])
def test_tile_source_chunks_follow_the_reference_loop(S, chunks):
    """acceleration_structures.py:76-113 steps source_end by floor(S / chunks) and
    sets it to S after the second-to-last chunk; a literal run of that loop."""
    from lyceanem.models.acceleration_structures import source_chunks

    exp, start, end = [], 0, int(np.floor(S / chunks))
    for i in range(chunks):
        exp.append((start, end))
        start = end
        end = int(end + np.floor(S / chunks))
        if i == chunks - 2:
            end = S
    assert source_chunks(S, chunks) == exp
    assert exp[-1][1] == S


def test_workchunking_three_waves_by_enumeration(oracle):
    """oracle.workchunking(..., 3) (rayfunctions.py:2863-2953) holds exactly the
    paths s -> k, s -> p -> k and s -> p -> q -> k (q != p) whose every leg is
    unblocked, with the reference's index values and zero padding; without the
    line of sight the first group is absent."""
    rng = np.random.default_rng(3)
    S, K, P = 3, 4, 6
    src = (rng.uniform(-1, 1, (S, 3)) * [1, 1, 0]).astype(np.float32)
    snk = (rng.uniform(-1, 1, (K, 3)) * [1, 1, 0] + [0, 0, 6]).astype(np.float32)
    sct = (rng.uniform(-2, 2, (P, 3)) * [1, 1, 0.5] + [0, 0, 3]).astype(np.float32)
    ctr = rng.uniform(-2, 2, (30, 1, 3)) * [1, 1, 0.6] + [0, 0, 3]
    tris = (ctr + rng.uniform(-0.5, 0.5, (30, 3, 3))).reshape(-1, 9).astype(np.float32)
    pts = np.concatenate([src, snk, sct])
    blk, _ = oracle.visibility_dense(pts, pts, tris, any_hit_only=True, want_dist=False)
    si, ki, pi = range(S), range(S, S + K), range(S + K, S + K + P)
    ok = lambda a, b: not blk[a, b]  # noqa: E731
    los = {(s + 1, k + 1, 0, 0) for s in si for k in ki if ok(s, k)}
    one = {(s + 1, p + 1, k + 1, 0) for s in si for p in pi for k in ki if ok(s, p) and ok(p, k)}
    two = {(s + 1, p + 1, q + 1, k + 1) for s in si for p in pi for q in pi for k in ki
           if q != p and ok(s, p) and ok(p, q) and ok(q, k)}
    assert los and one and two and 0.1 < blk.mean() < 0.9
    full, rays = oracle.workchunking(src, snk, sct, tris, 3)
    assert full.shape[1] == 4 and full.dtype == np.int32
    got = [tuple(int(x) for x in r) for r in full]
    assert len(got) == len(set(got)) and set(got) == los | one | two
    # rows come as line of sight, one bounce, two bounces (one source chunk)
    depth = (full[:, 1:] != 0).sum(1)
    assert np.all(np.diff(depth) >= 0)
    n_vis1 = sum(ok(s, p) for s in si for p in pi)
    n_vis2 = sum(ok(s, p) and ok(p, q) for s in si for p in pi for q in pi if q != p)
    assert rays == S * (K + P) + n_vis1 * (K + P - 1) + n_vis2 * K
    full_n, _ = oracle.workchunking(src, snk, sct, tris, 3, line_of_sight=False)
    assert set(tuple(int(x) for x in r) for r in full_n) == one | two


def test_scatter_branch_project_vectors_raises_like_the_reference():
    """calculate_scattering(scattering > 0, project_vectors=True) calls
    calculate_conformalVectors with two of its three arguments in the
    reference (frequency_domain.py:456-459, empropagation.py:3807) and raises
    TypeError; so does the drop-in, before any device work."""
    from lyceanem.electromagnetics import empropagation as EM
    from lyceanem.mesh import Mesh
    from lyceanem.models import frequency_domain as FD

    ap = Mesh(points=np.zeros((2, 3)), point_data={"Normals": np.tile([0, 0, 1.0], (2, 1)),
                                                   "Area": np.ones(2)})
    sk = Mesh(points=np.ones((3, 3)), point_data={"Normals": np.tile([0, 0, 1.0], (3, 1))})
    sc = Mesh(points=np.ones((2, 3)) * 2, point_data={"Normals": np.tile([0, 0, 1.0], (2, 1))})
    solid = Mesh(points=np.eye(3) + 5, cells=[("triangle", np.array([[0, 1, 2]]))])
    with pytest.raises(TypeError, match="antenna_axes"):
        FD.calculate_scattering(ap, sk, solid, np.array([[1.0, 0, 0]]), scatter_points=sc,
                                scattering=1, project_vectors=True)
    with pytest.raises(TypeError):
        EM.calculate_conformalVectors(np.array([1.0, 0, 0]), np.tile([0, 0, 1.0], (2, 1)))
