"""GPU parity: the HIP kernels through the C-ABI against the CPU oracle.

Bar: hit masks and kernel1Dv3 distances bit-exact; fields within 1e-4 relative
of the float64 oracle (norm-wise over sinks, and per sink against the sum of
term magnitudes), with the phase error stated.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

FIELD_RTOL = 1e-4


def _bits_to_bool(native, bits, K):
    return native.unpack_bits(bits, K)


def _c1_small(n=16, n_tris=2000, n_el=31, n_az=61, seed=1):
    from lyceanem import scenes

    return scenes.planar_array_c1(n=n, n_tris=n_tris, n_el=n_el, n_az=n_az, seed=seed)


@pytest.mark.parametrize("any_hit", [True, False])
def test_visibility_dense_exact(native, oracle, any_hit):
    sc = _c1_small()
    S, K = len(sc["src"]), len(sc["snk"])
    print(f"visibility_dense S={S} K={K} T={len(sc['tris'])} any_hit={any_hit}", flush=True)
    bits, dist = native.visibility_dense(sc["src"], sc["snk"], sc["tris"], any_hit=any_hit,
                                         want_dist=not any_hit)
    got = _bits_to_bool(native, bits, K)
    hit, odist = oracle.visibility_dense(sc["src"], sc["snk"], sc["tris"])
    assert 0.05 < hit.mean() < 0.95
    assert np.array_equal(got, hit), f"{(got != hit).sum()} mask mismatches"
    if not any_hit:
        d = dist.cpu().numpy()
        assert np.array_equal(d, odist), f"{(d != odist).sum()} dist mismatches"


def test_visibility_full_c1_sampled(native, oracle):
    from lyceanem import scenes

    sc = scenes.planar_array_c1()
    S, K = len(sc["src"]), len(sc["snk"])
    print(f"visibility_dense full C1 S={S} K={K}", flush=True)
    bits, _ = native.visibility_dense(sc["src"], sc["snk"], sc["tris"], any_hit=True)
    got = _bits_to_bool(native, bits, K)
    rng = np.random.default_rng(0)
    pairs = np.stack([rng.integers(0, S, 40000), rng.integers(0, K, 40000)], 1)
    hit, _ = oracle.visibility_sampled(sc["src"], sc["snk"], pairs, sc["tris"], any_hit_only=True)
    assert np.array_equal(got[pairs[:, 0], pairs[:, 1]], hit)


def test_visibility_edge_shapes(native, oracle):
    rng = np.random.default_rng(3)
    tris = (rng.uniform(0, 2, (37, 1, 3)) + rng.uniform(-0.5, 0.5, (37, 3, 3))).reshape(
        37, 9).astype(np.float32)
    for S, K in [(1, 1), (3, 65), (2, 130), (5, 257)]:
        src = rng.uniform(-1, 0, (S, 3)).astype(np.float32)
        snk = rng.uniform(2, 3, (K, 3)).astype(np.float32)
        bits, dist = native.visibility_dense(src, snk, tris, any_hit=False, want_dist=True)
        hit, odist = oracle.visibility_dense(src, snk, tris)
        assert np.array_equal(_bits_to_bool(native, bits, K), hit)
        assert np.array_equal(dist.cpu().numpy(), odist)
        # padding bits of the last word stay clear
        words = bits.cpu().numpy().view(np.uint64)
        if K % 64:
            assert (words[:, -1] >> np.uint64(K % 64) == 0).all()
    # no triangles: nothing is blocked
    bits, _ = native.visibility_dense(src, snk, np.empty((0, 9), np.float32))
    assert not _bits_to_bool(native, bits, K).any()


def test_native_semantics_and_skip_self(native, oracle):
    sc = _c1_small(n=8, n_tris=500, n_el=9, n_az=15)
    K = len(sc["snk"])
    bits, _ = native.visibility_dense(sc["src"], sc["snk"], sc["tris"], semantics=native.SEM_NATIVE)
    hit, _ = oracle.visibility_dense(sc["src"], sc["snk"], sc["tris"], semantics=1,
                                     any_hit_only=True, want_dist=False)
    assert np.array_equal(_bits_to_bool(native, bits, K), hit)
    bits, _ = native.visibility_dense(sc["src"], sc["snk"], sc["tris"], skip_self=True)
    hit, _ = oracle.visibility_dense(sc["src"], sc["snk"], sc["tris"], any_hit_only=True,
                                     skip_self=True, want_dist=False)
    assert np.array_equal(_bits_to_bool(native, bits, K), hit)


def test_raycast_pairs_exact(native, oracle):
    rng = np.random.default_rng(7)
    pts = rng.uniform(-2, 2, (300, 3)).astype(np.float32)
    ctr = rng.uniform(-1.5, 1.5, (700, 1, 3))
    tris = (ctr + rng.uniform(-0.3, 0.3, (700, 3, 3))).reshape(-1, 9).astype(np.float32)
    pairs = rng.integers(0, 300, (5000, 2)).astype(np.int32)
    pairs = pairs[pairs[:, 0] != pairs[:, 1]]
    inter, dist = native.raycast_pairs(pts, pairs, tris, any_hit=False, want_dist=True)
    d, L = oracle.charge(pts[pairs[:, 0]], pts[pairs[:, 1]])
    rays = np.concatenate([pts[pairs[:, 0]], d, L[:, None]], 1)
    oi, od = oracle.raycast_rays(rays, tris)
    assert np.array_equal(inter.cpu().numpy().astype(bool), oi)
    assert np.array_equal(dist.cpu().numpy(), od)
    inter2, _ = native.raycast_pairs(pts, pairs, tris, any_hit=True, want_dist=False)
    assert np.array_equal(inter2.cpu().numpy().astype(bool), oi)
    assert 0.05 < oi.mean() < 0.95


def _field_errors(got, ref, terms_scale):
    """(norm-wise relative error, max per-sink error / sum of |terms|)."""
    nrm = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    per = (np.abs(got - ref).max(axis=-1) / terms_scale).max()
    return nrm, per


def test_field_los_sum_vs_fp64_oracle(native, oracle):
    sc = _c1_small(n=16, n_tris=400, n_el=31, n_az=61)
    S, K = len(sc["src"]), len(sc["snk"])
    hit, _ = oracle.visibility_dense(sc["src"], sc["snk"], sc["tris"], any_hit_only=True,
                                     want_dist=False)
    bits, _ = native.visibility_dense(sc["src"], sc["snk"], sc["tris"])
    assert np.array_equal(_bits_to_bool(native, bits, K), hit)
    w = (sc["w"] * np.exp(1j * np.linspace(0, 6, S))[:, None]).astype(np.complex64)
    E = native.field_los(sc["src"], w, sc["snk"], sc["snk_n"], hit_bits=bits, alpha=0.0,
                         beta=sc["beta"]).cpu().numpy()
    ref = oracle.field_los(sc["src"], w, sc["snk"], sc["snk_n"], 0.0, sc["beta"], hit=hit)
    El = oracle.field_los(sc["src"], w, sc["snk"], sc["snk_n"], 0.0, sc["beta"], hit=hit,
                          elements=True)
    scale = np.abs(El).sum(axis=0).max(axis=-1) + 1e-30
    nrm, per = _field_errors(E, ref, scale)
    print(f"field_los sum: normwise {nrm:.2e}, per-sink {per:.2e}", flush=True)
    assert nrm < FIELD_RTOL and per < FIELD_RTOL


@pytest.mark.parametrize("S,K", [(10, 5), (5000, 3), (700, 130), (64, 4100)])
def test_field_los_chunk_counts(native, oracle, S, K):
    """The source sum is split into chunks by the sink count (about 2^21
    (sink, chunk) threads, at least 64 sources a chunk): one chunk (S < 64),
    the 64-source cap with few sinks, ragged chunks and a single chunk with
    many sink blocks all match the float64 oracle, with and without a mask
    and with lossy propagation."""
    rng = np.random.default_rng(S + K)
    src = rng.uniform(-1, 1, (S, 3)).astype(np.float32)
    snk = (rng.normal(size=(K, 3)) * 50).astype(np.float32)
    nrm = rng.normal(size=(K, 3)).astype(np.float32)
    w = (rng.normal(size=(S, 3)) + 1j * rng.normal(size=(S, 3))).astype(np.complex64)
    hit = rng.random((S, K)) < 0.3
    W = (K + 63) // 64
    packed = np.zeros((S, W), np.uint64)
    for k in range(K):
        packed[:, k >> 6] |= hit[:, k].astype(np.uint64) << np.uint64(k & 63)
    import torch
    bits = torch.from_numpy(packed.view(np.int64)).cuda()
    for alpha, mask in ((0.0, False), (0.0, True), (0.3, True)):
        E = native.field_los(src, w, snk, nrm, hit_bits=bits if mask else None, alpha=alpha,
                             beta=2 * np.pi / 0.1).cpu().numpy()
        ref = oracle.field_los(src, w, snk, nrm, alpha, 2 * np.pi / 0.1,
                               hit=hit if mask else None)
        El = oracle.field_los(src, w, snk, nrm, alpha, 2 * np.pi / 0.1,
                              hit=hit if mask else None, elements=True)
        scale = np.abs(El).sum(axis=0).max(axis=-1) + 1e-30
        nrm_err, per = _field_errors(E, ref, scale)
        print(f"chunks S={S} K={K} alpha={alpha} mask={mask}: {nrm_err:.2e} {per:.2e}", flush=True)
        assert nrm_err < FIELD_RTOL and per < FIELD_RTOL


def test_field_los_elements_and_phase(native, oracle):
    sc = _c1_small(n=8, n_tris=300, n_el=21, n_az=41)
    S, K = len(sc["src"]), len(sc["snk"])
    bits, _ = native.visibility_dense(sc["src"], sc["snk"], sc["tris"])
    hit = _bits_to_bool(native, bits, K)
    rng = np.random.default_rng(2)
    w = (rng.normal(size=(S, 3)) + 1j * rng.normal(size=(S, 3))).astype(np.complex64)
    for tangent in (False, True):
        E = native.field_los(sc["src"], w, sc["snk"], sc["snk_n"], hit_bits=bits,
                             beta=sc["beta"], elements=True, sink_tangent=tangent).cpu().numpy()
        ref = oracle.field_los(sc["src"], w, sc["snk"], sc["snk_n"], 0.0, sc["beta"], hit=hit,
                               elements=True, sink_tangent=tangent)
        big = np.abs(ref) > 1e-3 * np.abs(ref).max()
        rel = np.abs(E[big] - ref[big]) / np.abs(ref[big])
        dph = np.abs(np.angle(E[big] / ref[big]))
        print(f"elements tangent={tangent}: max rel {rel.max():.2e}, max phase {dph.max():.2e} rad",
              flush=True)
        assert rel.max() < FIELD_RTOL and dph.max() < 1e-4
        assert (E[hit] == 0).all()


@pytest.mark.parametrize("extent", [0.25, 5.0])
def test_field_long_paths_phase(native, oracle, extent):
    """1-10 km paths at 5.8 GHz, sources within +-0.25 m and +-5 m: the float64
    path length keeps the phase error far below 1e-3 rad (a float32 path length
    would not)."""
    rng = np.random.default_rng(4)
    lam = 299792458.0 / 5.8e9
    beta = 2 * np.pi / lam
    src = rng.uniform(-extent, extent, (64, 3)).astype(np.float32)
    u = rng.normal(size=(300, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    snk = (u * rng.uniform(1e3, 1e4, (300, 1))).astype(np.float32)
    w = np.zeros((64, 3), np.complex64)
    w[:, 2] = 1.0
    E = native.field_los(src, w, snk, u.astype(np.float32), beta=beta,
                         elements=True).cpu().numpy()
    ref = oracle.field_los(src, w, snk, u.astype(np.float32), 0.0, beta, elements=True)
    big = np.abs(ref) > 1e-3 * np.abs(ref).max()
    dph = np.abs(np.angle(E[big] / ref[big]))
    ref32 = oracle.field_los(src, w, snk, u.astype(np.float32), 0.0, beta, elements=True,
                             r_f32=True)
    dph32 = np.abs(np.angle(ref32[big] / ref[big]))
    print(f"long paths: max phase err {dph.max():.2e} rad (float32-r reference: "
          f"{dph32.max():.2e})", flush=True)
    assert dph.max() < 1e-4
    assert np.abs(E[big] - ref[big]).max() / np.abs(ref[big]).max() < FIELD_RTOL


@pytest.mark.parametrize("r_lo,r_hi", [(1e5, 1e6), (3e6, 3e7)])
def test_field_phase_very_long_paths(native, oracle, r_lo, r_hi):
    """Paths of 100 km - 30,000 km at 10 GHz (beta r up to 6e9 rad, 4e9 quarter
    turns): the quarter-turn rounding and quadrant of the phase reduction hold
    far beyond 2^31 quarter turns.  Per element against the float64 oracle, on
    the same float32 positions."""
    rng = np.random.default_rng(int(r_lo))
    beta = 2 * np.pi / 0.03
    src = rng.uniform(-0.2, 0.2, (32, 3)).astype(np.float32)
    u = rng.normal(size=(200, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    snk = (u * rng.uniform(r_lo, r_hi, (200, 1))).astype(np.float32)
    w = (rng.normal(size=(32, 3)) + 1j * rng.normal(size=(32, 3))).astype(np.complex64)
    E = native.field_los(src, w, snk, u.astype(np.float32), beta=beta,
                         elements=True).cpu().numpy()
    ref = oracle.field_los(src, w, snk, u.astype(np.float32), 0.0, beta, elements=True)
    big = np.abs(ref) > 1e-3 * np.abs(ref).max()
    dph = np.abs(np.angle(E[big] / ref[big]))
    rel = np.abs(E[big] - ref[big]) / np.abs(ref[big])
    print(f"paths {r_lo:.0e}-{r_hi:.0e} m: max phase err {dph.max():.2e} rad, max rel "
          f"{rel.max():.2e}", flush=True)
    assert dph.max() < 1e-4 and rel.max() < FIELD_RTOL


def test_field_los_mixed_source_extents(native, oracle):
    """One launch over compact and spread source chunks: 2048 sources in a
    0.2 m cube followed by 2048 spread over 40 m, sinks from 30 m to 10 km at
    10 GHz, with a mask; per element (phase and magnitude) and summed."""
    rng = np.random.default_rng(21)
    beta = 2 * np.pi / 0.03
    src = np.concatenate([rng.uniform(-0.1, 0.1, (2048, 3)),
                          rng.uniform(-20, 20, (2048, 3))]).astype(np.float32)
    u = rng.normal(size=(3, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    snk = (u * np.array([[30.0], [700.0], [1e4]])).astype(np.float32)
    nrm = u.astype(np.float32)
    S, K = len(src), len(snk)
    w = (rng.normal(size=(S, 3)) + 1j * rng.normal(size=(S, 3))).astype(np.complex64)
    hit = rng.random((S, K)) < 0.2
    packed = np.zeros((S, 1), np.uint64)
    for k in range(K):
        packed[:, 0] |= hit[:, k].astype(np.uint64) << np.uint64(k)
    import torch
    bits = torch.from_numpy(packed.view(np.int64)).cuda()
    El = oracle.field_los(src, w, snk, nrm, 0.0, beta, hit=hit, elements=True)
    big = np.abs(El) > 1e-3 * np.abs(El).max(axis=(0, 2), keepdims=True)
    for elements in (True, False):
        E = native.field_los(src, w, snk, nrm, hit_bits=bits, beta=beta,
                             elements=elements).cpu().numpy()
        if elements:
            dph = np.abs(np.angle(E[big] / El[big]))
            rel = np.abs(E[big] - El[big]) / np.abs(El[big])
            print(f"mixed extents, elements: max phase {dph.max():.2e} rad, max rel "
                  f"{rel.max():.2e}", flush=True)
            assert dph.max() < 1e-4 and rel.max() < FIELD_RTOL
        else:
            ref = El.sum(axis=0)
            scale = np.abs(El).sum(axis=0).max(axis=-1) + 1e-30
            nrm_err, per = _field_errors(E, ref, scale)
            print(f"mixed extents, sum: {nrm_err:.2e} {per:.2e}", flush=True)
            assert nrm_err < FIELD_RTOL and per < FIELD_RTOL


@pytest.mark.parametrize("ncol,magnetic", [(2, False), (3, False), (2, True), (3, True)])
def test_field_paths_vs_oracle(native, oracle, ncol, magnetic):
    """EMGPUFreqDomain on 2- and 3-column paths; with magnetic=True every other
    source is a magnetic point (Electric False, free-space impedance), whose
    paths launch (w x d) / eta."""
    rng = np.random.default_rng(10 + ncol)
    S, K, P = 6, 9, 7
    pos = np.concatenate([rng.uniform(-1, 1, (S, 3)), rng.uniform(20, 40, (K, 3)),
                          rng.uniform(3, 8, (P, 3))]).astype(np.float32)
    nrm = rng.normal(size=(S + K + P, 3)).astype(np.float32)
    w = (rng.normal(size=(S + K + P, 3)) + 1j * rng.normal(size=(S + K + P, 3))).astype(
        np.complex64)
    w[S:S + K] = 1.0
    los = np.array([[s + 1, S + 1 + k] for s in range(S) for k in range(K)
                    if rng.random() < 0.7])
    if ncol == 2:
        idx = los
    else:
        sc = np.array([[s + 1, S + K + 1 + p, S + 1 + k] for s in range(S) for p in range(P)
                       for k in range(K) if rng.random() < 0.5])
        idx = np.concatenate([np.concatenate([los, np.zeros((len(los), 1), int)], 1), sc])
    from lyceanem.electromagnetics import empropagation as EM
    from lyceanem.base_types import scattering_t

    table = np.zeros(S + K + P, scattering_t)
    for i, c in enumerate("xyz"):
        table["p" + c] = pos[:, i]
        table["n" + c] = nrm[:, i]
        table["e" + c] = w[:, i]
    table["Electric"] = True
    table["permittivity"] = 8.8541878187e-12
    table["permeability"] = 1.25663706212e-6
    eta = None
    if magnetic:
        table["Electric"][0:S:2] = False
        eta0 = np.float32(np.sqrt(np.float64(np.float32(1.25663706212e-6))
                                  / np.float64(np.float32(8.8541878187e-12))))
        eta = np.where(~table["Electric"], eta0, 0.0)
    net = EM.EMGPUFreqDomain(S, K, idx.astype(np.int32), table, 1.0, 0.0, 2 * np.pi / 0.5)
    ref = oracle.field_paths(pos, nrm, w, idx, S, K, 0.0, 2 * np.pi / 0.5, eta=eta)
    scale = np.abs(ref).max()
    assert np.abs(net - ref).max() < FIELD_RTOL * scale
    assert net.dtype == np.complex128 and net.shape == (S, K, 3)


@pytest.mark.parametrize("area_mode", ["per_source", "zero"])
def test_visiblespace_matches_oracle(native, oracle, area_mode):
    """rayfunctions.visiblespace: GPU visibility + projected-area kernel vs the
    oracle's visibility and the reference's angle / portion / patternsort."""
    from lyceanem.mesh import Mesh
    from lyceanem.raycasting import rayfunctions as RF
    from lyceanem.geometry.targets import spherical_field

    print(f"visiblespace {area_mode}", flush=True)
    rng = np.random.default_rng(21)
    g = (np.arange(8) - 3.5) * 0.1
    X, Y = np.meshgrid(g, g)
    src = np.stack([X.ravel(), Y.ravel(), np.full(64, 1e-4)], 1).astype(np.float32)
    nrm = np.tile(np.array([0, 0, 1.0], np.float32), (64, 1))
    tris = (np.array([0, 0, 0.6]) + rng.uniform(-0.5, 0.5, (30, 3, 3))).astype(np.float32)
    env = Mesh(points=tris.reshape(-1, 3), cells=[("triangle", np.arange(90).reshape(30, 3))])
    area = rng.uniform(0.005, 0.01, 64) if area_mode == "per_source" else 0
    az, el = np.linspace(-180, 180, 37), np.linspace(-90, 90, 19)
    from lyceanem.geometry import geometryfunctions as GF

    pat, shell = RF.visiblespace(Mesh(points=src), nrm, GF.mesh_conversion(env),
                                 vertex_area=area, az_range=az,
                                 elev_range=el, shell_range=3.0)
    snk = np.asarray(spherical_field(az, el, outward_normals=True, field_radius=3.0).points,
                     np.float32)
    hit, _ = oracle.visibility_dense(src, snk, tris.reshape(-1, 9), any_hit_only=True,
                                     want_dist=False)
    assert 0.05 < hit.mean() < 0.95
    ref = oracle.projected_area(src, nrm, area, snk, hit).reshape(len(el), len(az))
    assert pat.shape == ref.shape and pat.dtype == np.float32
    tol = 1e-5 * np.abs(ref).max()
    bad = np.abs(pat - ref) > tol
    if area_mode == "zero":
        # unit portions: a point whose dot product with the normal is within
        # float32 rounding of 0 may count on one side and not the other
        assert bad.sum() <= 2, np.argwhere(bad)
    else:
        np.testing.assert_allclose(pat, ref, rtol=1e-5, atol=tol)
