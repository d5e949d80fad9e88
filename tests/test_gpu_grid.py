"""GPU tests of the uniform-grid build and DDA traversal.

The grid must reproduce brute-force masks exactly (same predicate) and the
kernel1Dv3 distance (the closest accepted hit; ties closer than 2e-7 could
differ in principle, none occur in these scenes).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _small_tri_scene(seed, n_tris=3000, S=200, K=300, size=0.3):
    rng = np.random.default_rng(seed)
    ctr = rng.uniform(-2, 2, (n_tris, 1, 3))
    tris = (ctr + rng.uniform(-size, size, (n_tris, 3, 3))).reshape(-1, 9).astype(np.float32)
    src = rng.uniform(-3, 3, (S, 3)).astype(np.float32)
    snk = rng.uniform(-3, 3, (K, 3)).astype(np.float32)
    return src, snk, tris


@pytest.mark.parametrize("seed", [0, 1])
def test_grid_matches_brute_force(native, seed):
    src, snk, tris = _small_tri_scene(seed)
    K = len(snk)
    print(f"grid vs brute seed={seed}", flush=True)
    g = native.Grid(tris)
    assert g.info.nrefs >= len(tris)
    bits_g, dist_g = g.visibility(src, snk, any_hit=False, want_dist=True)
    bits_b, dist_b = native.visibility_dense(src, snk, tris, any_hit=False, want_dist=True)
    mg, mb = native.unpack_bits(bits_g, K), native.unpack_bits(bits_b, K)
    assert 0.05 < mb.mean() < 0.95
    assert np.array_equal(mg, mb), f"{(mg != mb).sum()} mask mismatches"
    dg, db = dist_g.cpu().numpy(), dist_b.cpu().numpy()
    assert np.array_equal(dg, db), f"{(dg != db).sum()} dist mismatches"
    bits_a, _ = g.visibility(src, snk, any_hit=True)
    assert np.array_equal(native.unpack_bits(bits_a, K), mb)


@pytest.mark.parametrize("sink_major", [False, True])
def test_grid_spatial_order_matches_input_order(native, sink_major):
    """The default traversal (lanes over a Morton order of the lane-side
    points, bits set atomically) gives the same masks and distances as lanes in
    input order, including ragged sizes, skip_self and a bits buffer that
    holds stale words."""
    src, snk, tris = _small_tri_scene(4, S=333, K=201)
    pts = np.concatenate([src[:150], snk[:100]])  # shared rows for skip_self
    g = native.Grid(tris)
    for any_hit in (False, True):
        for skip_self in (False, True):
            print(f"order sm={sink_major} any={any_hit} skip={skip_self}", flush=True)
            a_src, a_snk = (pts, pts[::-1].copy()) if skip_self else (src, snk)
            ref = g.visibility(a_src, a_snk, any_hit=any_hit, want_dist=not any_hit,
                               sink_major=sink_major, skip_self=skip_self, input_order=True)
            out = g.visibility(a_src, a_snk, any_hit=any_hit, want_dist=not any_hit,
                               sink_major=sink_major, skip_self=skip_self)
            assert np.array_equal(ref[0].cpu().numpy(), out[0].cpu().numpy())
            if not any_hit:
                assert np.array_equal(ref[1].cpu().numpy(), out[1].cpu().numpy())
            else:
                # walked from the sink end: the same mask
                for io in (False, True):
                    rev = g.visibility(a_src, a_snk, any_hit=True, sink_major=sink_major,
                                       skip_self=skip_self, input_order=io, from_sink=True)
                    assert np.array_equal(ref[0].cpu().numpy(), rev[0].cpu().numpy())
    # stale words in the caller's bits buffer are cleared by the call
    import torch
    lib = native.load_library()
    S, K = len(src), len(snk)
    rows, cols = (K, S) if sink_major else (S, K)
    flags = native.VIS_ANY_HIT | (native.VIS_SINK_MAJOR if sink_major else 0)
    bits = torch.full((rows, (cols + 63) // 64), -1, dtype=torch.int64, device="cuda")
    ws = g.vis_workspace(S, K, flags)
    s_t, k_t = native.to_dev(src, torch.float32), native.to_dev(snk, torch.float32)
    native.check(lib.lem_visibility_grid(native.ptr(s_t), S, native.ptr(k_t), K,
                                         native.ctypes.byref(g.info), native.ptr(g.cell_start),
                                         native.ptr(g.recs), native.ptr(g.cell_box),
                                         native.SEM_NUMBA, flags, native.ptr(bits), None,
                                         native.ptr(ws), ws.numel(), native.stream_ptr()), "vis")
    ref, _ = g.visibility(src, snk, any_hit=True, sink_major=sink_major, input_order=True)
    assert np.array_equal(bits.cpu().numpy(), ref.cpu().numpy())
    with pytest.raises(native.LemError, match="any-hit"):
        g.visibility(src, snk, any_hit=False, from_sink=True)


def test_grid_from_sink_matches_brute_force(native):
    """Rays ending just behind a triangle (a scatterer lifted off the far side
    of its face): the walk from the sink end finds the same masks as brute
    force, native semantics included."""
    src, _, tris = _small_tri_scene(9, n_tris=400, S=260, K=1)
    t = tris.reshape(-1, 3, 3).astype(np.float64)
    n = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    c = t.mean(1)
    pick = np.arange(0, len(t), 15)
    scat = np.concatenate([c[pick] + 1e-4 * n[pick], c[pick] - 1e-4 * n[pick]]).astype(np.float32)
    g = native.Grid(tris)
    for sem in (native.SEM_NUMBA, native.SEM_NATIVE):
        ref, _ = native.visibility_dense(src, scat, tris, semantics=sem, any_hit=True)
        m = native.unpack_bits(ref, len(scat))
        assert 0.05 < m.mean() < 0.95
        for sm in (False, True):
            b, _ = g.visibility(src, scat, any_hit=True, sink_major=sm, semantics=sem,
                                from_sink=True)
            mb = native.unpack_bits(b, len(src)).T if sm else native.unpack_bits(b, len(scat))
            assert np.array_equal(mb, m), f"sem={sem} sm={sm}: {(mb != m).sum()} mismatches"


@pytest.mark.parametrize("S", [524288, 524289])
def test_grid_mask_rows_at_the_lds_limit(native, S):
    """Sink-major rows of 8,192 words (64 KB: lane-order words, reordered in
    LDS) and of 8,193 words (per-ray atomic ORs into cleared rows), any-hit
    and closest-hit, against the brute-force kernel; two sinks, one of them
    also a source (row boundaries off the word grid)."""
    src, snk, tris = _small_tri_scene(12, n_tris=500, S=S, K=2)
    snk[1] = src[S // 3]
    g = native.Grid(tris)
    print(f"grid rows of {S} lane-side points", flush=True)
    ref, ref_d = native.visibility_dense(src, snk, tris, any_hit=False, want_dist=True)
    mb = native.unpack_bits(ref, len(snk))
    assert 0.05 < mb.mean() < 0.95
    for any_hit in (True, False):
        b, d = g.visibility(src, snk, any_hit=any_hit, sink_major=True, want_dist=not any_hit)
        m = native.unpack_bits(b, S).T
        assert np.array_equal(m, mb), f"any_hit={any_hit}: {(m != mb).sum()} mismatches"
        if not any_hit:
            assert np.array_equal(d.cpu().numpy().T, ref_d.cpu().numpy())


def test_grid_claims_interleave_rows(native):
    """Closest-hit walks over rows long enough (400,000 lane-side points: 66
    claims per XCD share) claim them in interleaved groups of 64 rows; 100
    sinks give one full group and a last group of 36 (any-hit walks take one
    row at a time).  Masks and distances equal the brute-force kernel."""
    S, K = 400_000, 100
    src, snk, tris = _small_tri_scene(21, n_tris=500, S=S, K=K)
    g = native.Grid(tris)
    print(f"interleaved claims: {S} x {K}", flush=True)
    ref, ref_d = native.visibility_dense(src, snk, tris, any_hit=False, want_dist=True)
    mb = native.unpack_bits(ref, K)
    assert 0.05 < mb.mean() < 0.95
    for any_hit in (True, False):
        b, d = g.visibility(src, snk, any_hit=any_hit, sink_major=True, want_dist=not any_hit)
        m = native.unpack_bits(b, S).T
        assert np.array_equal(m, mb), f"any_hit={any_hit}: {(m != mb).sum()} mismatches"
        if not any_hit:
            assert np.array_equal(d.cpu().numpy().T, ref_d.cpu().numpy())


def test_grid_sink_major_layout(native):
    src, snk, tris = _small_tri_scene(3, S=130, K=70)
    g = native.Grid(tris)
    bits, dist = g.visibility(src, snk, any_hit=False, want_dist=True)
    bt, dt = g.visibility(src, snk, any_hit=False, want_dist=True, sink_major=True)
    m = native.unpack_bits(bits, len(snk))
    mt = native.unpack_bits(bt, len(src))
    assert np.array_equal(m, mt.T)
    assert np.array_equal(dist.cpu().numpy(), dt.cpu().numpy().T)


def test_grid_axis_aligned_and_outside_rays(native, oracle):
    """Rays along the axes, along cell planes, starting and ending outside the
    grid, and through a closed box."""
    rng = np.random.default_rng(5)
    # a cube of 12 triangles plus a floor of 2 triangles
    v = np.array([[0, 0, 0], [1, 0, 0], [1, 1, 0], [0, 1, 0], [0, 0, 1], [1, 0, 1], [1, 1, 1],
                  [0, 1, 1]], np.float32)
    f = [(0, 1, 2), (0, 2, 3), (4, 5, 6), (4, 6, 7), (0, 1, 5), (0, 5, 4), (1, 2, 6), (1, 6, 5),
         (2, 3, 7), (2, 7, 6), (3, 0, 4), (3, 4, 7)]
    cube = np.array([np.concatenate([v[a], v[b], v[c]]) for a, b, c in f], np.float32)
    floor = np.array([[-5, -5, -1, 5, -5, -1, 5, 5, -1], [-5, -5, -1, 5, 5, -1, -5, 5, -1]],
                     np.float32)
    clutter = (rng.uniform(-4, 4, (300, 1, 3)) * [1, 1, 0.2] + [0, 0, 3] +
               rng.uniform(-0.3, 0.3, (300, 3, 3))).reshape(-1, 9).astype(np.float32)
    tris = np.concatenate([cube, floor, clutter])
    pts = [[0.5, 0.5, -3.0], [0.5, 0.5, 5.0], [-6.0, 0.5, 0.5], [6.0, 0.5, 0.5],
           [0.5, -6.0, 0.5], [0.5, 6.0, 0.5], [0.0, 0.0, 9.0], [1.0, 1.0, -9.0],
           [2.0, 2.0, 2.0], [-2.0, -2.0, -0.5], [0.5, 0.5, 0.5]]
    src = np.array(pts, np.float32)
    snk = np.array(pts[::-1], np.float32)
    g = native.Grid(tris)
    bits_g, dist_g = g.visibility(src, snk, any_hit=False, want_dist=True)
    hit, odist = oracle.visibility_dense(src, snk, tris)
    assert np.array_equal(native.unpack_bits(bits_g, len(snk)), hit)
    assert np.array_equal(dist_g.cpu().numpy(), odist)


@pytest.mark.parametrize("n_par", ["some", "overflow"])
def test_grid_rays_along_grid_planes(native, n_par):
    """Rays with a zero direction component are queued by the persistent
    traversal and walked by a second kernel; past the queue's capacity
    (262,144 rays) that kernel rechecks every ray.  Masks, distances and the
    any-hit, sink-major and from-sink variants equal the brute force."""
    rng = np.random.default_rng(11)
    ctr = rng.uniform(-2, 2, (1500, 1, 3))
    tris = (ctr + rng.uniform(-0.3, 0.3, (1500, 3, 3))).reshape(-1, 9).astype(np.float32)
    if n_par == "some":
        # sources on the plane z = 0.25; a third of the sinks on it too, and
        # some sinks sharing a source's x exactly
        S, K = 300, 400
        src = rng.uniform(-3, 3, (S, 3)).astype(np.float32)
        src[:, 2] = 0.25
        snk = rng.uniform(-3, 3, (K, 3)).astype(np.float32)
        snk[::3, 2] = 0.25
        snk[1::7, 0] = src[: len(snk[1::7]), 0]
    else:
        # every ray in the plane z = 0.25: 640 x 480 = 307,200 parallel rays
        S, K = 640, 480
        src = rng.uniform(-3, 3, (S, 3)).astype(np.float32)
        snk = rng.uniform(-3, 3, (K, 3)).astype(np.float32)
        src[:, 2] = 0.25
        snk[:, 2] = 0.25
    print(f"rays along grid planes: {n_par}, S={S} K={K}", flush=True)
    g = native.Grid(tris)
    assert g.info.nmacro <= 4352  # the persistent traversal runs
    bits_b, dist_b = native.visibility_dense(src, snk, tris, any_hit=False, want_dist=True)
    mb, db = native.unpack_bits(bits_b, K), dist_b.cpu().numpy()
    assert 0.05 < mb.mean() < 0.95
    for sm in (False, True):
        bits, dist = g.visibility(src, snk, any_hit=False, want_dist=True, sink_major=sm)
        m = native.unpack_bits(bits, S).T if sm else native.unpack_bits(bits, K)
        d = dist.cpu().numpy().T if sm else dist.cpu().numpy()
        assert np.array_equal(m, mb), f"sink_major={sm}: {(m != mb).sum()} mask mismatches"
        assert np.array_equal(d, db), f"sink_major={sm}: {(d != db).sum()} dist mismatches"
        bits_a, _ = g.visibility(src, snk, any_hit=True, sink_major=sm)
        ma = native.unpack_bits(bits_a, S).T if sm else native.unpack_bits(bits_a, K)
        assert np.array_equal(ma, mb)
        for io in (False, True):
            bits_i, _ = g.visibility(src, snk, any_hit=True, sink_major=sm, input_order=io)
            mi = native.unpack_bits(bits_i, S).T if sm else native.unpack_bits(bits_i, K)
            assert np.array_equal(mi, mb), f"sink_major={sm} input_order={io}"
    bits_f, _ = g.visibility(src, snk, any_hit=True, sink_major=True, from_sink=True)
    assert np.array_equal(native.unpack_bits(bits_f, S).T, mb)


def test_grid_terrain_small_vs_oracle(native, oracle):
    from lyceanem import scenes

    sc = scenes.terrain_c2(nqx=100, nqy=50, size_x=100.0, size_y=50.0, n_src_side=32,
                           n_snk=256, corr=8.0, src_alt=10.0)
    src, snk, tris = sc["src"], sc["snk"], sc["tris"]
    g = native.Grid(tris)
    bits, dist = g.visibility(src, snk, any_hit=False, want_dist=True, sink_major=True)
    m = native.unpack_bits(bits, len(src)).T
    d = dist.cpu().numpy().T
    hit, odist = oracle.visibility_dense(src, snk, tris)
    print(f"terrain small: blocked fraction {hit.mean():.3f}", flush=True)
    assert 0.02 < hit.mean() < 0.98
    assert np.array_equal(m, hit), f"{(m != hit).sum()} mask mismatches"
    assert np.array_equal(d, odist), f"{(d != odist).sum()} dist mismatches"


def test_grid_flat_terrain_vs_oracle(native, oracle):
    """A flat terrain (few cells vertically, most rays grazing the surface):
    every ray's mask and closest-hit distance equal the brute force, for both
    lane layouts, and the any-hit mask walked from the sink end too."""
    from lyceanem import scenes

    sc = scenes.terrain_c2(nqx=200, nqy=100, size_x=200.0, size_y=100.0, n_src_side=16,
                           n_snk=128, corr=8.0, sigma_h=1.0, src_alt=6.0)
    src, snk, tris = sc["src"], sc["snk"], sc["tris"]
    g = native.Grid(tris)
    print(f"flat terrain: {tuple(g.info.n)} cells", flush=True)
    hit, odist = oracle.visibility_dense(src, snk, tris)
    assert 0.02 < hit.mean() < 0.98
    for sm in (False, True):
        bits, dist = g.visibility(src, snk, any_hit=False, want_dist=True, sink_major=sm)
        m = native.unpack_bits(bits, len(src)).T if sm else native.unpack_bits(bits, len(snk))
        d = dist.cpu().numpy().T if sm else dist.cpu().numpy()
        assert np.array_equal(m, hit), f"sink_major={sm}: {(m != hit).sum()} mask mismatches"
        assert np.array_equal(d, odist), f"sink_major={sm}: {(d != odist).sum()} dist mismatches"
    bits_f, _ = g.visibility(src, snk, any_hit=True, sink_major=True, from_sink=True)
    assert np.array_equal(native.unpack_bits(bits_f, len(src)).T, hit)


def _sample_bits(bits, rows, cols):
    """Bits (rows[i], cols[i]) of a (rows, words) int64 device tensor -> bool numpy."""
    import torch

    r = torch.from_numpy(rows).to(bits.device)
    c = torch.from_numpy(cols).to(bits.device)
    w = bits[r, c >> 6]
    return ((w >> (c & 63)) & 1).bool().cpu().numpy()


def test_grid_terrain_full_launch_sampled(native, oracle):
    """The benchmarked launch itself: all 1,048,576 sources x 16,384 sinks of
    the 1,000,000-triangle terrain (sink-major, closest hit, mask + distance).
    1,000,000 random rays of it are compared with the brute-force kernel over
    every triangle (lem_raycast_pairs, itself bit-exact against the oracle),
    and 10,000 of them with the oracle directly; mismatch counts are printed."""
    import torch
    from lyceanem import scenes

    sc = scenes.terrain_c2()
    src, snk, tris = sc["src"], sc["snk"], sc["tris"]
    S, K = len(src), len(snk)
    assert tris.shape == (1000000, 9)
    g = native.Grid(tris)
    print(f"terrain grid: {tuple(g.info.n)} cells, {g.info.nrefs} refs; full launch "
          f"{S} x {K}", flush=True)
    bits, dist = g.visibility(src, snk, any_hit=False, want_dist=True, sink_major=True)
    rng = np.random.default_rng(1)
    n = 1_000_000
    s_i, k_i = rng.integers(0, S, n), rng.integers(0, K, n)
    m_grid = _sample_bits(bits, k_i, s_i)
    d_grid = dist[torch.from_numpy(k_i).to(dist.device),
                  torch.from_numpy(s_i).to(dist.device)].cpu().numpy()
    # any-hit mask of the same launch
    del dist
    bits_a, _ = g.visibility(src, snk, any_hit=True, sink_major=True)
    m_any = _sample_bits(bits_a, k_i, s_i)
    del bits, bits_a
    torch.cuda.empty_cache()
    print("terrain full launch: brute force on the sample", flush=True)
    pts = np.concatenate([src, snk])
    pairs = np.stack([s_i, S + k_i], 1).astype(np.int32)
    inter, bd = native.raycast_pairs(pts, pairs, tris, any_hit=False, want_dist=True)
    m_b, d_b = inter.cpu().numpy().astype(bool), bd.cpu().numpy()
    mm, dm, am = int((m_grid != m_b).sum()), int((d_grid != d_b).sum()), \
        int((m_any != m_b).sum())
    print(f"terrain full launch: {n} rays vs brute force: {mm} mask, {dm} distance, "
          f"{am} any-hit mask mismatches (blocked {m_b.mean():.4f})", flush=True)
    no = 10_000
    hit, odist = oracle.visibility_sampled(src, snk, np.stack([s_i[:no], k_i[:no]], 1), tris)
    om, od = int((m_grid[:no] != hit).sum()), int((d_grid[:no] != odist).sum())
    print(f"terrain full launch: {no} rays vs the oracle: {om} mask, {od} distance "
          f"mismatches", flush=True)
    assert 0.02 < m_b.mean() < 0.98
    assert (mm, dm, am, om, od) == (0, 0, 0, 0, 0)


def test_fused_grid_field_matches_masked_field(native):
    src, snk, tris = _small_tri_scene(7, S=150, K=230)
    rng = np.random.default_rng(7)
    w = (rng.normal(size=(len(src), 3)) + 1j * rng.normal(size=(len(src), 3))).astype(np.complex64)
    nrm = rng.normal(size=(len(snk), 3)).astype(np.float32)
    wk = (rng.normal(size=(len(snk), 3)) + 1j * rng.normal(size=(len(snk), 3))).astype(np.complex64)
    g = native.Grid(tris)
    bits, _ = g.visibility(src, snk, any_hit=True)
    for tangent in (False, True):
        a = native.field_los(src, w, snk, nrm, hit_bits=bits, snk_w=wk, beta=2 * np.pi / 0.3,
                             sink_tangent=tangent).cpu().numpy()
        b = native.field_los_grid(g, src, w, snk, nrm, snk_w=wk, beta=2 * np.pi / 0.3,
                                  sink_tangent=tangent).cpu().numpy()
        # identical masks and per-pair arithmetic, same summation order
        assert np.array_equal(a, b)
        # lanes over the sinks in input order instead of the Morton order
        c = native.field_los_grid(g, src, w, snk, nrm, snk_w=wk, beta=2 * np.pi / 0.3,
                                  sink_tangent=tangent, input_order=True).cpu().numpy()
        assert np.array_equal(a, c)


def test_sink_major_mask_field_matches_fused(native, term_bound):
    """The field read through a sink-major mask (K x W64(S)), whole or in sink
    chunks, equals the fused grid field and the source-major masked field."""
    src, snk, tris = _small_tri_scene(8, S=150, K=230)
    rng = np.random.default_rng(8)
    w = (rng.normal(size=(len(src), 3)) + 1j * rng.normal(size=(len(src), 3))).astype(np.complex64)
    nrm = rng.normal(size=(len(snk), 3)).astype(np.float32)
    wk = (rng.normal(size=(len(snk), 3)) + 1j * rng.normal(size=(len(snk), 3))).astype(np.complex64)
    g = native.Grid(tris)
    beta = 2 * np.pi / 0.3
    bits_sm, _ = g.visibility(src, snk, any_hit=True, sink_major=True)
    bits, _ = g.visibility(src, snk, any_hit=True)
    m = native.unpack_bits(bits, len(snk))
    assert np.array_equal(native.unpack_bits(bits_sm, len(src)).T, m)
    assert 0.05 < m.mean() < 0.95
    for tangent in (False, True):
        ref = native.field_los_grid(g, src, w, snk, nrm, snk_w=wk, beta=beta,
                                    sink_tangent=tangent).cpu().numpy()
        a = native.field_los(src, w, snk, nrm, hit_bits=bits_sm, snk_w=wk, beta=beta,
                             sink_tangent=tangent, bits_sink_major=True).cpu().numpy()
        assert np.array_equal(a, ref)
        # 5 sink chunks of 50 rows (one row is 3 words here)
        for cb in (1 << 31, 24 * 50):
            b = native.field_los_grid_sink_major(g, src, w, snk, nrm, snk_w=wk, beta=beta,
                                                 sink_tangent=tangent,
                                                 chunk_bytes=cb).cpu().numpy()
            assert np.array_equal(b, ref)
    # source-major mask from the persistent traversal: whole (identical sums)
    # and in source chunks (chunk sums added: float32 summation rounding, below
    # 1e-5 of the sum of the sink's |terms|)
    for tangent in (False, True):
        ref = native.field_los_grid(g, src, w, snk, nrm, snk_w=wk, beta=beta,
                                    sink_tangent=tangent).cpu().numpy()
        a = native.field_los_grid_masked(g, src, w, snk, nrm, snk_w=wk, beta=beta,
                                         sink_tangent=tangent, sink_major=False).cpu().numpy()
        assert np.array_equal(a, ref)
        b = native.field_los_grid_masked(g, src, w, snk, nrm, snk_w=wk, beta=beta,
                                         sink_tangent=tangent, sink_major=False,
                                         chunk_bytes=32 * 40).cpu().numpy()
        bound = term_bound(src, w, snk, nrm, wk, 0.0, beta)
        per = (np.abs(b - ref).max(axis=-1) / bound).max()
        print(f"source-chunked masked field: per-sink {per:.2e}", flush=True)
        assert per < 1e-5
    # per-element network from either layout
    e1 = native.field_los(src, w, snk, nrm, hit_bits=bits, snk_w=wk, beta=beta,
                          elements=True).cpu().numpy()
    e2 = native.field_los(src, w, snk, nrm, hit_bits=bits_sm, snk_w=wk, beta=beta,
                          elements=True, bits_sink_major=True).cpu().numpy()
    assert np.array_equal(e1, e2)
    assert np.array_equal(e1[m.astype(bool)], np.zeros_like(e1[m.astype(bool)]))
    with pytest.raises(native.LemError, match="hit_bits shape"):
        native.field_los(src, w, snk, nrm, hit_bits=bits, bits_sink_major=True)


@pytest.mark.parametrize("S,K", [(1, 1), (1, 65), (65, 1), (3, 130), (130, 3)])
def test_grid_tiny_and_ragged_shapes(native, S, K):
    """Shapes where rows hold fewer than 8 words (some XCD owners get none),
    single rays and ragged last words: every layout and mode equals brute force."""
    rng = np.random.default_rng(S * 1000 + K)
    ctr = rng.uniform(-2, 2, (2000, 1, 3))
    tris = (ctr + rng.uniform(-0.3, 0.3, (2000, 3, 3))).reshape(-1, 9).astype(np.float32)
    src = rng.uniform(-3, 3, (S, 3)).astype(np.float32)
    snk = rng.uniform(-3, 3, (K, 3)).astype(np.float32)
    g = native.Grid(tris)
    ref_b, ref_d = native.visibility_dense(src, snk, tris, any_hit=False, want_dist=True)
    m = native.unpack_bits(ref_b, K)
    for sm in (False, True):
        b, d = g.visibility(src, snk, any_hit=False, want_dist=True, sink_major=sm)
        mb = native.unpack_bits(b, S).T if sm else native.unpack_bits(b, K)
        db = d.cpu().numpy().T if sm else d.cpu().numpy()
        assert np.array_equal(mb, m)
        assert np.array_equal(db, ref_d.cpu().numpy())
        for fs in (False, True):
            a, _ = g.visibility(src, snk, any_hit=True, sink_major=sm, from_sink=fs)
            ma = native.unpack_bits(a, S).T if sm else native.unpack_bits(a, K)
            assert np.array_equal(ma, m)
