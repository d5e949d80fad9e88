"""GPU tests of the drop-in Python API against the oracle's pipeline
(workchunkingv2 -> freqdomainkernal restated on the CPU)."""
import numpy as np
import pytest

from lyceanem.mesh import Mesh

pytestmark = pytest.mark.gpu


def _scene(n=6, n_tris=60, seed=3):
    rng = np.random.default_rng(seed)
    c = (np.arange(n) - (n - 1) / 2) * 0.05
    xx, yy = np.meshgrid(c, c)
    src = np.stack([xx.ravel(), yy.ravel(), np.zeros(n * n)], 1)
    ctr = rng.uniform(-1, 1, (n_tris, 1, 3)) * [1, 1, 0.3] + [0, 0, 0.8]
    tv = (ctr + rng.uniform(-0.2, 0.2, (n_tris, 3, 3))).reshape(-1, 3)
    solid = Mesh(points=tv, cells=[("triangle", np.arange(3 * n_tris).reshape(-1, 3))])
    aperture = Mesh(points=src, cells=[],
                    point_data={"Normals": np.tile([0.0, 0, 1], (n * n, 1)),
                                "Area": np.full(n * n, 0.0025)})
    return aperture, solid


def _sinks(n_el=9, n_az=13, r=50.0):
    from lyceanem.geometry.targets import spherical_field

    return spherical_field(np.linspace(-180, 180, n_az), np.linspace(-90, 90, n_el),
                           outward_normals=True, field_radius=r)


def test_workchunking_matches_oracle(oracle, monkeypatch):
    from lyceanem.raycasting import rayfunctions as RF
    from lyceanem.geometry import geometryfunctions as GF
    from lyceanem.base_types import triangles_as_array

    aperture, solid = _scene()
    snk = _sinks().points.astype(np.float32)
    src = aperture.points.astype(np.float32)
    tris = GF.mesh_conversion(solid)
    rng = np.random.default_rng(1)
    scat = rng.uniform(-1, 1, (11, 3)).astype(np.float32) * [1, 1, 0.2] + [0, 0, 1.4]
    for ms in (1, 2):
        got, rays = RF.workchunkingv2(src, snk, scat.astype(np.float32), tris, ms)
        exp, erays = oracle.workchunking(src, snk, scat, triangles_as_array(tris), ms)
        assert np.array_equal(got, exp) and rays == erays
        assert got.shape[0] > 0
    # chunked == unchunked when free memory is tiny
    monkeypatch.setattr(RF, "free_memory_bytes", lambda: 40000)
    got1, _ = RF.workchunkingv2(src, snk, np.empty((0, 3), np.float32), tris, 1)
    exp1, _ = oracle.workchunking(src, snk, np.empty((0, 3), np.float32),
                                  triangles_as_array(tris), 1)
    assert np.array_equal(got1, exp1)
    got2, _ = RF.workchunkingv2(src, snk, scat.astype(np.float32), tris, 2)
    exp2, _ = oracle.workchunking(src, snk, scat, triangles_as_array(tris), 2)
    assert sorted(map(tuple, got2)) == sorted(map(tuple, exp2))


@pytest.mark.parametrize("elements", [False, True])
def test_calculate_scattering_los(oracle, elements):
    from lyceanem.models import frequency_domain as FD
    from lyceanem.geometry import geometryfunctions as GF
    from lyceanem.base_types import triangles_as_array

    aperture, solid = _scene()
    sinks = _sinks()
    lam = 0.1
    desired = np.array([[1.0, 0.5j, 0.0]])
    Ex, Ey, Ez = FD.calculate_scattering(aperture, sinks, solid, desired, wavelength=lam,
                                         elements=elements, beta=2 * np.pi / lam)
    src = aperture.points.astype(np.float32)
    snk = sinks.points.astype(np.float32)
    tris = triangles_as_array(GF.mesh_conversion(solid))
    hit, _ = oracle.visibility_dense(src, snk, tris, any_hit_only=True, want_dist=False)
    w = (np.repeat(desired.astype(np.complex64), len(src), 0) * 0.0025).astype(np.complex64)
    ref = oracle.field_los(src, w, snk, sinks.point_data["Normals"], 0.0, 2 * np.pi / lam,
                           hit=hit, elements=elements)
    got = np.stack([Ex, Ey, Ez], -1)
    assert got.dtype == np.complex128 and got.shape == ref.shape
    assert 0 < hit.mean() < 1
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-4


def test_calculate_scattering_one_bounce(oracle):
    from lyceanem.models import frequency_domain as FD
    from lyceanem.geometry import geometryfunctions as GF
    from lyceanem.base_types import triangles_as_array

    aperture, solid = _scene(n=4, n_tris=30)
    sinks = _sinks(5, 9)
    rng = np.random.default_rng(9)
    sp = rng.uniform(-1, 1, (13, 3)) * [1, 1, 0.1] + [0, 0, 1.5]
    scat = Mesh(points=sp, cells=[], point_data={"Normals": np.tile([0.0, 0, -1], (13, 1))})
    lam = 0.2
    Ex, Ey, Ez = FD.calculate_scattering(aperture, sinks, solid, np.array([[0.0, 1.0, 0.0]]),
                                         scatter_points=scat, wavelength=lam, scattering=1,
                                         beta=2 * np.pi / lam)
    src = aperture.points.astype(np.float32)
    snk = sinks.points.astype(np.float32)
    tris = triangles_as_array(GF.mesh_conversion(solid))
    full, _ = oracle.workchunking(src, snk, sp.astype(np.float32), tris, 2)
    S, K = len(src), len(snk)
    pos = np.concatenate([src, snk, sp]).astype(np.float32)
    nrm = np.concatenate([aperture.point_data["Normals"], sinks.point_data["Normals"],
                          scat.point_data["Normals"]]).astype(np.float32)
    w = np.ones((len(pos), 3), np.complex64)
    w[:S] = np.array([0, 1, 0], np.complex64)
    net = oracle.field_paths(pos, nrm, w, full, S, K, 0.0, 2 * np.pi / lam)
    ref = net.sum(0)
    got = np.stack([Ex, Ey, Ez], -1)
    assert (full[:, 2] != 0).any()
    assert np.linalg.norm(got - ref) / np.linalg.norm(ref) < 1e-4


def test_calculate_farfield_unit_sphere_quirk(oracle):
    from lyceanem.models import frequency_domain as FD

    aperture, solid = _scene(n=4, n_tris=5)
    az = np.linspace(-180, 180, 13)
    el = np.linspace(-90, 90, 7)
    et, ep = FD.calculate_farfield(aperture, solid, np.array([[1.0, 0, 0]]), az_range=az,
                                   el_range=el, wavelength=0.1, farfield_distance=20,
                                   beta=2 * np.pi / 0.1)
    assert et.shape == (7 * 13,) and ep.shape == (7 * 13,)
    from lyceanem.geometry.targets import spherical_field

    sinks = spherical_field(az, el, True, 1.0)  # radius 1, outward, whatever farfield_distance
    Ex, Ey, Ez = FD.calculate_scattering(aperture, sinks, solid, np.array([[1.0, 0, 0]]),
                                         wavelength=0.1, beta=2 * np.pi / 0.1)
    th, ph = sinks.point_data["theta_(Radians)"], sinks.point_data["phi_(Radians)"]
    et2, ep2 = oracle.etheta_ephi(Ex, Ey, Ez, th, ph)
    assert np.allclose(et, et2) and np.allclose(ep, ep2)


def test_one_bounce_grid_equals_brute_force_masks(oracle):
    """one_bounce_fields with visibility through the grid (fused kernel) equals
    the masked brute-force legs."""
    from lyceanem import _native
    from lyceanem.models.frequency_domain import one_bounce_fields

    rng = np.random.default_rng(21)
    ctr = rng.uniform(-3, 3, (900, 1, 3)) * [1, 1, 0.3] + [0, 0, 4]
    tris = (ctr + rng.uniform(-0.4, 0.4, (900, 3, 3))).reshape(-1, 9).astype(np.float32)
    src = rng.uniform(-0.5, 0.5, (60, 3)).astype(np.float32) * [1, 1, 0]
    w = (rng.normal(size=(60, 3)) + 1j * rng.normal(size=(60, 3))).astype(np.complex64)
    scat = (rng.uniform(-3, 3, (150, 3)) * [1, 1, 0.3] + [0, 0, 4]).astype(np.float32)
    scat_n = rng.normal(size=(150, 3)).astype(np.float32)
    snk = (rng.uniform(-6, 6, (90, 3)) * [1, 1, 0] + [0, 0, 9]).astype(np.float32)
    snk_n = np.tile(np.array([[0, 0, 1]], np.float32), (90, 1))
    beta = 2 * np.pi / 0.3
    a = one_bounce_fields(src, w, snk, snk_n, scat, scat_n, tris, beta=beta)
    b = one_bounce_fields(src, w, snk, snk_n, scat, scat_n, tris, beta=beta,
                          grid=_native.Grid(tris))
    assert np.abs(a).max() > 0
    assert np.allclose(a, b, rtol=1e-6, atol=1e-6 * np.abs(a).max())


@pytest.mark.parametrize("use_grid", [False, True])
def test_one_bounce_no_visible_scatterer_rule(oracle, use_grid):
    """workchunkingv2 returns no rows at all, line of sight included, when every
    source -> scatterer ray is blocked (rayfunctions.py:2835-2861).  The decision
    comes from the leg-1 mask: scatterers that are visible but carry zero weight
    keep the line of sight."""
    from lyceanem import _native
    from lyceanem.models.frequency_domain import one_bounce_fields

    rng = np.random.default_rng(5)
    # a lid at z = 1 over the sources; scatterers above it, sinks off to the side
    lid = np.array([[-5, -5, 1, 5, -5, 1, 0, 5, 1]], np.float32)
    clutter = (rng.uniform(-3, 3, (40, 1, 3)) * [1, 1, 0.1] + [0, 0, 4] +
               rng.uniform(-0.2, 0.2, (40, 3, 3))).reshape(-1, 9).astype(np.float32)
    tris = np.concatenate([lid, clutter])
    src = (rng.uniform(-0.3, 0.3, (20, 3)) * [1, 1, 0]).astype(np.float32)
    w = (rng.normal(size=(20, 3)) + 1j * rng.normal(size=(20, 3))).astype(np.complex64)
    scat = (rng.uniform(-2, 2, (30, 3)) * [1, 1, 0] + [0, 0, 2]).astype(np.float32)
    scat_n = np.tile(np.array([[0, 0, -1]], np.float32), (30, 1))
    snk = (rng.uniform(-1, 1, (25, 3)) * [1, 1, 0] + [40, 0, 0.5]).astype(np.float32)
    snk_n = np.tile(np.array([[-1, 0, 0]], np.float32), (25, 1))
    beta = 2 * np.pi / 0.3
    grid = _native.Grid(tris) if use_grid else None
    full, _ = oracle.workchunking(src, snk, scat, tris, 2)
    assert full.shape[0] == 0  # the lid blocks every source -> scatterer ray
    E = one_bounce_fields(src, w, snk, snk_n, scat, scat_n, tris, beta=beta, grid=grid)
    assert np.array_equal(E, np.zeros_like(E))
    # no lid: scatterers visible, but with zero weight every H_p is zero; the
    # line-of-sight rows stay (the oracle's path sum)
    tris2 = clutter
    grid2 = _native.Grid(tris2) if use_grid else None
    full2, _ = oracle.workchunking(src, snk, scat, tris2, 2)
    assert (full2[:, 2] != 0).any() and (full2[:, 2] == 0).any()
    zero = np.zeros((30, 3), np.complex64)
    E2 = one_bounce_fields(src, w, snk, snk_n, scat, scat_n, tris2, scat_w=zero, beta=beta,
                           grid=grid2)
    S, K = len(src), len(snk)
    pos = np.concatenate([src, snk, scat])
    nrm = np.concatenate([np.tile([[0, 0, 1.0]], (S, 1)), snk_n, scat_n]).astype(np.float32)
    wt = np.concatenate([w, np.ones((K, 3), np.complex64), zero])
    ref = oracle.field_paths(pos, nrm, wt, full2, S, K, 0.0, beta).sum(0)
    assert np.abs(ref).max() > 0
    assert np.linalg.norm(E2 - ref) / np.linalg.norm(ref) < 1e-5


def _x_vertical_terrain(n=10, seed=2):
    """Height field with x vertical (the orientation the tile grid bins in y, z)."""
    rng = np.random.default_rng(seed)
    y, z = np.meshgrid(np.linspace(-1, 1, n + 1), np.linspace(-1, 1, n + 1))
    x = 0.15 * rng.normal(size=y.shape)
    pts = np.stack([x.ravel(), y.ravel(), z.ravel()], 1)
    idx = np.arange((n + 1) ** 2).reshape(n + 1, n + 1)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, :-1].ravel(), \
        idx[1:, 1:].ravel()
    tri = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)])
    tri = tri[rng.random(len(tri)) < 0.35]
    return Mesh(points=pts, cells=[("triangle", tri)])


@pytest.mark.parametrize("scattering", [0, 1])
def test_calculate_scattering_cuda_branch_vs_oracle(oracle, scattering):
    """calculate_scattering(cuda=True) (frequency_domain.py:719-926) with
    calculate_scattering_cuda (:195-264): native culled predicate, source row ==
    sink row pairs untested, weights x Area; one bounce as the source ->
    scatterer network, the summed scatterer weights and the matrix product of
    the two networks per component, against the oracle."""
    from lyceanem.models import frequency_domain as FD

    solid = _x_vertical_terrain()
    tris9 = np.asarray(solid.points, np.float32)[solid.cells[0].data].reshape(-1, 9)
    rng = np.random.default_rng(8)
    S, K, P = 12, 30, 9
    src = np.stack([rng.uniform(0.5, 1.5, S), rng.uniform(-0.8, 0.8, S),
                    rng.uniform(-0.8, 0.8, S)], 1)
    snk = np.stack([rng.uniform(-1.5, -0.5, K), rng.uniform(-0.8, 0.8, K),
                    rng.uniform(-0.8, 0.8, K)], 1)
    scp = np.stack([rng.uniform(-0.4, 0.4, P), rng.uniform(-0.8, 0.8, P),
                    rng.uniform(-0.8, 0.8, P)], 1)
    area = rng.uniform(0.5, 1.5, S)
    aperture = Mesh(points=src, point_data={"Normals": np.tile([-1.0, 0, 0], (S, 1)),
                                            "Area": area})
    sinks = Mesh(points=snk, point_data={"Normals": np.tile([1.0, 0, 0], (K, 1))})
    scat = Mesh(points=scp, point_data={"Normals": rng.normal(size=(P, 3))})
    desired = np.array([[0.2, 1.0, 0.5j]])
    lam = 0.2
    beta = 2 * np.pi / lam
    out = {}
    for elements in (False, True):
        Ex, Ey, Ez = FD.calculate_scattering(aperture, sinks, solid, desired,
                                             scatter_points=scat, wavelength=lam,
                                             scattering=scattering, elements=elements,
                                             beta=beta, cuda=True)
        out[elements] = np.stack([Ex, Ey, Ez], -1)
    src32, snk32, scp32 = (x.astype(np.float32) for x in (src, snk, scp))
    w = (np.repeat(desired.astype(np.complex64), S, 0) * area[:, None]).astype(np.complex64)

    def native_net(a, wa, b, b_n):
        hit, _ = oracle.visibility_dense(a, b, tris9, semantics=1, any_hit_only=True,
                                         skip_self=True, want_dist=False)
        return oracle.field_los(a, wa, b, b_n, 0.0, beta, hit=hit, elements=True), hit

    net, hit = native_net(src32, w, snk32, sinks.point_data["Normals"])
    assert 0.05 < hit.mean() < 0.95
    if scattering == 1:
        src_scat, _ = native_net(src32, w, scp32, scat.point_data["Normals"])
        ws = src_scat.sum(0).astype(np.complex64)
        scat_snk, _ = native_net(scp32, ws, snk32, sinks.point_data["Normals"])
        net = net + np.einsum("spc,pkc->skc", src_scat, scat_snk)
    for elements, ref in ((True, net), (False, net.sum(0))):
        err = np.linalg.norm(out[elements] - ref) / np.linalg.norm(ref)
        print(f"cuda branch scattering={scattering} elements={elements}: {err:.2e}", flush=True)
        assert out[elements].shape == ref.shape and err < 1e-4


def test_aperture_projection_vs_oracle(oracle):
    """aperture_projection (frequency_domain.py:15-91): triangle centroids lifted
    by 1e-7 along their normals, weighted by the aperture's own cell areas, the
    visible projected area on the shell as 4 pi A / lambda^2; the caller's mesh
    is left unchanged, and a mesh without cell areas raises as the
    reference's does."""
    from lyceanem.geometry import geometryfunctions as GF
    from lyceanem.geometry.targets import spherical_field
    from lyceanem.models import frequency_domain as FD

    g = np.linspace(-0.5, 0.5, 5)
    X, Y = np.meshgrid(g, g)
    pts = np.stack([X.ravel(), Y.ravel(), np.zeros(25)], 1)
    idx = np.arange(25).reshape(5, 5)
    a, b, c, d = idx[:-1, :-1].ravel(), idx[:-1, 1:].ravel(), idx[1:, :-1].ravel(), \
        idx[1:, 1:].ravel()
    tri = np.concatenate([np.stack([a, b, d], 1), np.stack([a, d, c], 1)])
    plate = GF.compute_areas(Mesh(points=pts, cells=[("triangle", tri)]))
    lid = Mesh(points=np.array([[-0.2, -2, 1.0], [0.9, -2, 1.0], [0.3, 2, 1.0]]),
               cells=[("triangle", np.array([[0, 1, 2]]))])
    az = np.linspace(-180, 180, 13)
    el = np.linspace(-90, 90, 9)
    lam = 0.3
    before = {k: np.array(v[0], copy=True) for k, v in plate.cell_data.items()}
    env, pcd = FD.aperture_projection(plate, environment=[plate, lid], wavelength=lam,
                                      az_range=az, elev_range=el, farfield_distance=5.0)
    assert set(plate.cell_data) == set(before)
    for k, v in before.items():
        assert np.array_equal(plate.cell_data[k][0], v)
    cen = (pts[tri].mean(1) + 1e-7 * np.array([0, 0, 1.0])).astype(np.float32)
    shell = spherical_field(az, el, outward_normals=True, field_radius=5.0)
    tris9 = np.concatenate([pts[tri].reshape(-1, 9), lid.points.reshape(1, 9)]).astype(
        np.float32)
    hit, _ = oracle.visibility_dense(cen, shell.points, tris9, any_hit_only=True,
                                     want_dist=False)
    assert 0.01 < hit.mean() < 0.99
    nrm = np.tile(np.array([[0, 0, 1.0]], np.float32), (len(tri), 1))
    pat = oracle.projected_area(cen, nrm, plate.cell_data["Area"][0], shell.points, hit)
    ref = 4 * np.pi * pat.reshape(len(el), len(az)) / lam ** 2
    assert env.shape == (len(el), len(az)) and env.dtype == np.float32
    assert np.allclose(env, ref, rtol=1e-5, atol=1e-6 * ref.max())
    assert np.allclose(pcd.point_data["Directivity_Envelope"].reshape(env.shape), env, rtol=1e-5)
    bare = Mesh(points=pts, cells=[("triangle", tri)])
    with pytest.raises(KeyError, match="Area"):
        FD.aperture_projection(bare, wavelength=lam, az_range=az, elev_range=el)
