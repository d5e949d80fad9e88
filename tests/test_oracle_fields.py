"""Pins the oracle's float64 field restatement (freqdomainkernal et al.).

Independent checks: the textbook Huygens far-field kernel j/(lambda r) e^{-jkr}
cos(theta) that the reference's Rayleigh-Sommerfeld factor tends to for
r >> lambda, the closed-form array factor of a steered uniform planar array,
the projector identity of the launch basis, linearity, masking, and the
two-leg factorisation the GPU one-bounce path relies on.
"""
import numpy as np
import pytest

LAM = 0.1
BETA = 2 * np.pi / LAM


def test_single_pair_huygens_far_field(oracle):
    rng = np.random.default_rng(1)
    src = np.zeros((1, 3), np.float32)
    snk = (rng.normal(size=(20, 3)) * 1000).astype(np.float32)
    r = np.linalg.norm(snk.astype(np.float64), axis=1)
    u = snk / r[:, None]
    nrm = u.astype(np.float32)  # sink normal along the path: cos(theta) = 1
    w = np.array([[0.3 + 0.1j, -0.2j, 0.7]], np.complex64)
    E = oracle.field_los(src, w, snk, nrm, 0.0, BETA)
    w64 = w[0].astype(np.complex128)
    perp = w64[None, :] - u * (u @ w64)[:, None]
    huy = (1j / LAM) * np.exp(-1j * BETA * r) / r
    expect = huy[:, None] * perp
    # the -(1/r) term of dG is 1/(beta r) ~ 1.6e-5 relative at 1 km
    assert np.allclose(E, expect, rtol=5e-5, atol=1e-12 * np.abs(expect).max())
    # the obliquity factor is the raw d . n (negative when the normal faces away)
    E2 = oracle.field_los(src, w, snk, -nrm, 0.0, BETA)
    assert np.allclose(E2, -E, rtol=1e-12)


def test_attenuation_alpha(oracle):
    src = np.zeros((1, 3), np.float32)
    snk = np.array([[0, 0, 50.0], [0, 0, 100.0]], np.float32)
    nrm = np.array([[0, 0, 1], [0, 0, 1]], np.float32)
    w = np.array([[1, 0, 0]], np.complex64)
    E0 = oracle.field_los(src, w, snk, nrm, 0.0, BETA)
    Ea = oracle.field_los(src, w, snk, nrm, 0.01, BETA)
    ratio = np.abs(Ea[:, 0]) / np.abs(E0[:, 0])
    # |(-(a+jb) - 1/r)| / |(-jb - 1/r)| * e^{-a r}
    corr = np.abs(0.01 + 1j * BETA + 1 / np.array([50, 100])) / np.abs(
        1j * BETA + 1 / np.array([50, 100]))
    assert np.allclose(ratio, np.exp(-0.01 * np.array([50.0, 100.0])) * corr, rtol=1e-10)


def test_steered_planar_array_factor(oracle):
    """Sum over an N x N lambda/2 array steered to u0 equals the closed-form
    array factor times the element term in the far field."""
    n = 8
    pitch = LAM / 2
    c = (np.arange(n) - (n - 1) / 2) * pitch
    xx, yy = np.meshgrid(c, c)
    src = np.stack([xx.ravel(), yy.ravel(), np.zeros(n * n)], 1).astype(np.float32)
    th0, ph0 = np.radians(20.0), np.radians(35.0)
    u0 = np.array([np.sin(th0) * np.cos(ph0), np.sin(th0) * np.sin(ph0), np.cos(th0)])
    # e^{-j beta r} propagation: r_s ~ R - p_s.u, so weights e^{-j beta p_s.u0} steer to u0
    phase = np.exp(-1j * BETA * (src.astype(np.float64) @ u0))
    w = np.zeros((n * n, 3), np.complex64)
    w[:, 0] = phase
    R = 1.0e6
    th = np.radians(np.array([0.0, 10.0, 20.0, 20.0, 40.0]))
    ph = np.radians(np.array([0.0, 35.0, 35.0, 100.0, 200.0]))
    uu = np.stack([np.sin(th) * np.cos(ph), np.sin(th) * np.sin(ph), np.cos(th)], 1)
    snk = (uu * R).astype(np.float32)
    E = oracle.field_los(src, w, snk, uu.astype(np.float32), 0.0, BETA)
    psx = BETA * pitch * (uu[:, 0] - u0[0])
    psy = BETA * pitch * (uu[:, 1] - u0[1])

    def af1(p):
        with np.errstate(invalid="ignore", divide="ignore"):
            v = np.sin(n * p / 2) / np.sin(p / 2)
        return np.where(np.abs(np.sin(p / 2)) < 1e-12, n, v)

    AF = af1(psx) * af1(psy)
    r = np.linalg.norm(snk.astype(np.float64), axis=1)
    xhat = np.array([1.0, 0, 0])
    perp = xhat[None, :] - uu * uu[:, [0]]
    elem = (1j / LAM) * np.exp(-1j * BETA * r) / r
    expect = (AF * elem)[:, None] * perp
    scale = np.abs(expect).max()
    assert np.allclose(E, expect, rtol=1e-3, atol=2e-4 * scale)
    # main beam at u0 is the largest response of the grid
    assert np.argmax(np.abs(E[:, 0])) in (2,)


def test_projector_identity_and_tangent_quirk(oracle):
    rng = np.random.default_rng(2)
    src = rng.uniform(-1, 1, (5, 3)).astype(np.float32)
    snk = rng.uniform(4, 9, (6, 3)).astype(np.float32)
    nrm = rng.normal(size=(6, 3)).astype(np.float32)
    w = (rng.normal(size=(5, 3)) + 1j * rng.normal(size=(5, 3))).astype(np.complex64)
    E = oracle.field_los(src, w, snk, nrm, 0.0, BETA, elements=True)
    d = snk[None].astype(np.float64) - src[:, None]
    d /= np.linalg.norm(d, axis=2, keepdims=True)
    assert np.abs(np.einsum("skc,skc->sk", E, d)).max() < 1e-12 * np.abs(E).max()
    # launch basis (E.u)u + (E.v)v equals (I - dd^T)E: E is parallel to the projection
    proj = w[:, None, :] - d * np.einsum("sc,skc->sk", w.astype(np.complex128), d)[..., None]
    ratio = E / proj
    assert np.allclose(ratio, ratio[..., [0]], rtol=1e-9)
    Et = oracle.field_los(src, w, snk, nrm, 0.0, BETA, elements=True, sink_tangent=True)
    nh = nrm / np.linalg.norm(nrm.astype(np.float64), axis=1, keepdims=True)
    assert np.abs(np.einsum("skc,kc->sk", Et, nh)).max() < 1e-12 * np.abs(Et).max()
    assert np.allclose(Et, E - nh[None] * np.einsum("skc,kc->sk", E, nh)[..., None], atol=1e-14)


def test_linearity_and_mask(oracle):
    rng = np.random.default_rng(4)
    src = rng.uniform(-1, 1, (9, 3)).astype(np.float32)
    snk = rng.uniform(3, 9, (7, 3)).astype(np.float32)
    nrm = rng.normal(size=(7, 3)).astype(np.float32)
    w1 = (rng.normal(size=(9, 3)) + 1j * rng.normal(size=(9, 3))).astype(np.complex64)
    w2 = (rng.normal(size=(9, 3)) + 1j * rng.normal(size=(9, 3))).astype(np.complex64)
    a, b = 0.5 - 0.25j, -1.5 + 2j
    wc = (a * w1.astype(np.complex128) + b * w2).astype(np.complex64)
    E1 = oracle.field_los(src, w1, snk, nrm, 0.0, BETA)
    E2 = oracle.field_los(src, w2, snk, nrm, 0.0, BETA)
    Ec = oracle.field_los(src, wc, snk, nrm, 0.0, BETA)
    assert np.allclose(Ec, a * E1 + b * E2, rtol=1e-6, atol=1e-7 * np.abs(Ec).max())
    hit = rng.random((9, 7)) < 0.4
    Em = oracle.field_los(src, w1, snk, nrm, 0.0, BETA, hit=hit)
    El = oracle.field_los(src, w1, snk, nrm, 0.0, BETA, elements=True)
    assert np.allclose(Em, (El * (~hit)[..., None]).sum(0), rtol=1e-12, atol=1e-18)
    assert np.allclose(El.sum(0), E1, rtol=1e-12)


def test_paths_match_dense_and_two_leg_factorisation(oracle):
    rng = np.random.default_rng(6)
    S, K, P = 4, 5, 6
    src = rng.uniform(-1, 1, (S, 3)).astype(np.float32)
    snk = rng.uniform(5, 9, (K, 3)).astype(np.float32)
    sct = rng.uniform(1, 4, (P, 3)).astype(np.float32)
    src_n = np.tile(np.array([[0, 0, 1]], np.float32), (S, 1))
    snk_n = rng.normal(size=(K, 3)).astype(np.float32)
    sct_n = rng.normal(size=(P, 3)).astype(np.float32)
    ws = (rng.normal(size=(S, 3)) + 1j * rng.normal(size=(S, 3))).astype(np.complex64)
    wp = (0.5 * np.ones((P, 3))).astype(np.complex64)
    pos = np.concatenate([src, snk, sct])
    nrm = np.concatenate([src_n, snk_n, sct_n])
    wts = np.concatenate([ws, np.ones((K, 3), np.complex64), wp])
    # 2-column index == dense LOS network
    idx2 = np.array([[s + 1, S + 1 + k] for s in range(S) for k in range(K)], np.int64)
    net2 = oracle.field_paths(pos, nrm, wts, idx2, S, K, 0.0, BETA)
    El = oracle.field_los(src, ws, snk, snk_n, 0.0, BETA, elements=True)
    assert np.allclose(net2, El, rtol=1e-12, atol=1e-18)
    # 3-column index: LOS rows get the sink-tangent projection (reference quirk)
    los = np.concatenate([idx2, np.zeros((len(idx2), 1), np.int64)], 1)
    sc = np.array([[s + 1, S + K + 1 + p, S + 1 + k] for s in range(S) for p in range(P)
                   for k in range(K)], np.int64)
    net3 = oracle.field_paths(pos, nrm, wts, np.concatenate([los, sc]), S, K, 0.0, BETA)
    Elt = oracle.field_los(src, ws, snk, snk_n, 0.0, BETA, elements=True, sink_tangent=True)
    # two-leg factorisation of the scattered part, source-summed
    H = oracle.field_los(src, ws, sct, sct_n, 0.0, BETA, snk_w=wp, sink_tangent=True)
    Es = oracle.field_los(sct, H.astype(np.complex64), snk, snk_n, 0.0, BETA)
    total = net3.sum(0)
    # H is rounded to complex64 between the legs -> 1e-6 relative
    assert np.allclose(total, Elt.sum(0) + Es, rtol=2e-6, atol=2e-6 * np.abs(total).max())


def test_r_f32_error_is_reported_scale(oracle):
    """Rounding r to float32 as the numba kernel does moves the phase of a 1 km
    path by up to beta * ulp(1000)/2 ~ 2e-3 rad."""
    src = np.zeros((1, 3), np.float32)
    rng = np.random.default_rng(8)
    u = rng.normal(size=(200, 3))
    u /= np.linalg.norm(u, axis=1, keepdims=True)
    snk = (u * 1000).astype(np.float32)
    w = np.array([[1, 0, 0]], np.complex64)
    E = oracle.field_los(src, w, snk, u.astype(np.float32), 0.0, BETA)
    E32 = oracle.field_los(src, w, snk, u.astype(np.float32), 0.0, BETA, r_f32=True)
    ok = np.abs(E[:, 0]) > 1e-9
    dph = np.abs(np.angle(E32[ok, 0] / E[ok, 0]))
    assert 1e-4 < dph.max() < BETA * 6.2e-5 / 2 * 1.01


def test_torch64_restatement_matches_c_oracle(oracle):
    """The closed-form float64 torch field (used for reference gradients) equals
    the C oracle's u/v-basis restatement."""
    import torch

    rng = np.random.default_rng(12)
    S, K = 11, 17
    src = rng.uniform(-1, 1, (S, 3)).astype(np.float32)
    snk = rng.uniform(3, 9, (K, 3)).astype(np.float32)
    nrm = rng.normal(size=(K, 3)).astype(np.float32)
    w = (rng.normal(size=(S, 3)) + 1j * rng.normal(size=(S, 3))).astype(np.complex64)
    hit = rng.random((S, K)) < 0.3
    for alpha in (0.0, 0.02):
        ref = oracle.field_los(src, w, snk, nrm, alpha, BETA, hit=hit)
        t = oracle.field_los_torch64(torch.from_numpy(w.astype(np.complex128)),
                                     torch.from_numpy(src.astype(np.float64)),
                                     torch.from_numpy(snk.astype(np.float64)),
                                     torch.from_numpy(nrm.astype(np.float64)),
                                     torch.from_numpy(~hit), alpha, BETA).numpy()
        assert np.allclose(t, ref, rtol=1e-10, atol=1e-14 * np.abs(ref).max())


def test_projected_area_closed_form(oracle):
    """visiblespace's pattern of one unblocked source facing +z: area * cos(theta)
    in the upper half-space, 0 below; with vertex_area == 0 every visible point
    counts 1 (rayfunctions.py:405-417)."""
    rng = np.random.default_rng(3)
    d = rng.normal(size=(400, 3))
    d /= np.linalg.norm(d, axis=1, keepdims=True)
    snk = (2.0 * d).astype(np.float32)
    src = np.zeros((1, 3), np.float32)
    nrm = np.array([[0, 0, 1]], np.float32)
    free = np.zeros((1, 400), bool)
    pat = oracle.projected_area(src, nrm, np.array([0.25]), snk, free)
    np.testing.assert_allclose(pat, 0.25 * np.maximum(d[:, 2], 0), rtol=1e-5, atol=1e-6)
    ones = oracle.projected_area(src, nrm, 0, snk, free)
    np.testing.assert_array_equal(ones, (d[:, 2] > 1e-6).astype(np.float32) * (ones > 0))
    assert ones[d[:, 2] > 1e-3].min() == 1.0 and ones[d[:, 2] < -1e-3].max() == 0.0
    blocked = oracle.projected_area(src, nrm, np.array([0.25]), snk, np.ones((1, 400), bool))
    assert not blocked.any()
    # two sources add; per-source areas are used
    src2 = np.array([[0, 0, 0], [0.1, 0, 0]], np.float32)
    nrm2 = np.array([[0, 0, 1], [0, 0, 1]], np.float32)
    both = oracle.projected_area(src2, nrm2, np.array([0.25, 0.5]), snk, np.zeros((2, 400), bool))
    one = oracle.projected_area(src2[1:], nrm2[1:], np.array([0.5]), snk, np.zeros((1, 400), bool))
    np.testing.assert_allclose(both, pat + one, rtol=1e-5, atol=1e-6)


def _proj(E, n):
    n = np.asarray(n, np.float64)
    n = n / np.linalg.norm(n)
    return E - n * (E @ n)


def _leg(a, b, n_b, beta):
    """lossy_propagation's factor and direction of one leg a -> b (float64)."""
    D = np.asarray(b, np.float64) - np.asarray(a, np.float64)
    r = np.linalg.norm(D)
    d = D / r
    g = 1j * beta
    L = -(1 / (2 * np.pi)) * (-g - 1 / r) * np.exp(-g * r) / r * (d @ np.asarray(n_b, np.float64))
    return L, d


def test_four_column_rows_closed_form(oracle):
    """Rows of a 4-column index (workchunkingv2 with max_scatter 3) against the
    per-row closed form of freqdomainkernal (empropagation.py:1083-1265): a
    zero column at position i >= 1 reads point -1 (the last row of the table,
    numba's negative-index wrap), so a line-of-sight row [s, k, 0, 0] is
    projected onto the sink's tangent plane and then onto the last point's; a
    one-bounce row [s, p, k, 0] ends with the sink's tangent projection; a
    two-bounce row [s, p, q, k] has none at the sink."""
    rng = np.random.default_rng(12)
    pos = rng.uniform(-3, 3, (5, 3)).astype(np.float32)  # s, k, p, q, last
    nrm = rng.normal(size=(5, 3)).astype(np.float32)
    w = (rng.normal(size=(5, 3)) + 1j * rng.normal(size=(5, 3))).astype(np.complex64)
    w[1] = 1.0  # the sink weight (1 in calculate_scattering)
    P64, N64, W = pos.astype(np.float64), nrm.astype(np.float64), w.astype(np.complex128)
    S, K = 1, 1
    # line of sight
    net = oracle.field_paths(pos, nrm, w, np.array([[1, 2, 0, 0]]), S, K, 0.0, BETA)
    L, d = _leg(P64[0], P64[1], N64[1], BETA)
    expect = L * _proj(_proj(_proj(W[0], d) * W[1], N64[1]), N64[4])
    assert np.allclose(net[0, 0], expect, rtol=1e-12, atol=1e-15)
    # the last-point projection really acts (not a sink-only projection)
    only_sink = L * _proj(_proj(W[0], d) * W[1], N64[1])
    assert np.abs(expect - only_sink).max() > 1e-3 * np.abs(expect).max()
    # one bounce: s -> p -> k
    net = oracle.field_paths(pos, nrm, w, np.array([[1, 3, 2, 0]]), S, K, 0.0, BETA)
    L1, d1 = _leg(P64[0], P64[2], N64[2], BETA)
    L2, d2 = _leg(P64[2], P64[1], N64[1], BETA)
    e = _proj(_proj(W[0], d1) * W[2], N64[2])
    expect = L1 * L2 * _proj(_proj(e, d2) * W[1], N64[1])
    assert np.allclose(net[0, 0], expect, rtol=1e-12, atol=1e-15)
    # two bounces: s -> p -> q -> k
    net = oracle.field_paths(pos, nrm, w, np.array([[1, 3, 4, 2]]), S, K, 0.0, BETA)
    L3, d3 = _leg(P64[2], P64[3], N64[3], BETA)
    L4, d4 = _leg(P64[3], P64[1], N64[1], BETA)
    e = _proj(_proj(_proj(W[0], d1) * W[2], N64[2]), d3) * W[3]
    expect = L1 * L3 * L4 * (_proj(_proj(e, N64[3]), d4) * W[1])
    assert np.allclose(net[0, 0], expect, rtol=1e-12, atol=1e-15)


def test_four_column_paths_three_leg_factorisation(oracle):
    """Every path of a 4-column index summed over sources equals the three-leg
    evaluation the GPU two-bounce path uses (scatterer fields H, second-bounce
    fields G with the self pairs left out, the extra projections of the
    line-of-sight and one-bounce rows)."""
    rng = np.random.default_rng(8)
    S, K, P = 3, 4, 5
    src = rng.uniform(-1, 1, (S, 3)).astype(np.float32)
    snk = rng.uniform(5, 9, (K, 3)).astype(np.float32)
    sct = rng.uniform(1, 4, (P, 3)).astype(np.float32)
    src_n = np.tile(np.array([[0, 0, 1]], np.float32), (S, 1))
    snk_n = rng.normal(size=(K, 3)).astype(np.float32)
    sct_n = rng.normal(size=(P, 3)).astype(np.float32)
    ws = (rng.normal(size=(S, 3)) + 1j * rng.normal(size=(S, 3))).astype(np.complex64)
    wp = (0.5 * np.ones((P, 3))).astype(np.complex64)
    pos = np.concatenate([src, snk, sct])
    nrm = np.concatenate([src_n, snk_n, sct_n])
    wts = np.concatenate([ws, np.ones((K, 3), np.complex64), wp])
    s1, k1, p1 = np.arange(S) + 1, S + 1 + np.arange(K), S + K + 1 + np.arange(P)
    rows = [[s, k, 0, 0] for s in s1 for k in k1]
    rows += [[s, p, k, 0] for s in s1 for p in p1 for k in k1]
    rows += [[s, p, q, k] for s in s1 for p in p1 for q in p1 if q != p for k in k1]
    net = oracle.field_paths(pos, nrm, wts, np.array(rows, np.int64), S, K, 0.0, BETA)
    n_last = sct_n[-1].astype(np.float64)
    n_last = n_last / np.linalg.norm(n_last)
    Elt = oracle.field_los(src, ws, snk, snk_n, 0.0, BETA, sink_tangent=True)
    E0 = Elt - n_last[None, :] * (Elt @ n_last)[:, None]
    H = oracle.field_los(src, ws, sct, sct_n, 0.0, BETA, snk_w=wp, sink_tangent=True)
    H32 = H.astype(np.complex64)
    E1 = oracle.field_los(sct, H32, snk, snk_n, 0.0, BETA, sink_tangent=True)
    G = oracle.field_los(sct, H32, sct, sct_n, 0.0, BETA, snk_w=wp, sink_tangent=True,
                         hit=np.eye(P, dtype=bool))
    E2 = oracle.field_los(sct, G.astype(np.complex64), snk, snk_n, 0.0, BETA)
    total = net.sum(0)
    # H and G are rounded to complex64 between the legs -> ~1e-6 relative
    assert np.allclose(total, E0 + E1 + E2, rtol=5e-6, atol=5e-6 * np.abs(total).max())


ETA0 = float(np.sqrt(1.25663706212e-6 / 8.8541878187e-12))


def test_magnetic_source_closed_form(oracle):
    """A path from a magnetic point launches (w x d) / eta instead of w
    (empropagation.py:1111-1146).  w = z with the sink on +x gives the field of
    an electric source of weight y / eta (z x x = y); for random complex w and
    random 2- and 3-point paths the field equals that of electric sources of
    weight numpy.cross(w, d) / eta, d the unit direction of the first leg."""
    # the closed-form case: z x x = +y (a transposed cross product gives -y)
    pos = np.array([[0, 0, 0], [200, 0, 0]], np.float32)
    nrm = np.array([[0, 0, 1], [1, 0, 0]], np.float32)
    w = np.array([[0, 0, 1], [1, 1, 1]], np.complex64)
    idx = np.array([[1, 2]])
    eta = np.array([ETA0, 0.0])
    Em = oracle.field_paths(pos, nrm, w, idx, 1, 1, 0.0, BETA, eta=eta)
    we = np.array([[0, 1 / ETA0, 0], [1, 1, 1]], np.complex64)
    Ee = oracle.field_paths(pos, nrm, we.astype(np.complex128), idx, 1, 1, 0.0, BETA)
    assert np.abs(Em[0, 0, 1]) > 0 and Em[0, 0, 0] == 0 and Em[0, 0, 2] == 0
    np.testing.assert_allclose(Em, Ee, rtol=1e-6)
    # random paths: S sources, K sinks, P scatterers, half the sources magnetic
    rng = np.random.default_rng(17)
    S, K, P = 8, 5, 6
    N = S + K + P
    pos = rng.uniform(-30, 30, (N, 3)).astype(np.float32)
    nrm = rng.normal(size=(N, 3)).astype(np.float32)
    w = (rng.normal(size=(N, 3)) + 1j * rng.normal(size=(N, 3))).astype(np.complex64)
    eta = np.where(np.arange(N) < S, np.where(np.arange(N) % 2 == 0, ETA0, 0.0), 0.0)
    rows = [[s + 1, S + k + 1, 0] for s in range(S) for k in range(K)]
    rows += [[s + 1, S + K + p + 1, S + k + 1] for s in range(S) for p in range(P)
             for k in range(K)]
    idx = np.array(rows)
    Em = oracle.field_paths(pos, nrm, w, idx, S, K, 0.0, BETA, eta=eta)
    Ee = np.zeros_like(Em)
    for row in idx:
        a, b = row[0] - 1, row[1] - 1
        d = pos[b].astype(np.float64) - pos[a]
        d /= np.linalg.norm(d)
        wr = w.astype(np.complex128)
        if eta[a] > 0:
            wr[a] = np.cross(wr[a], d) / eta[a]
        Ee += oracle.field_paths(pos, nrm, wr, row[None, :], S, K, 0.0, BETA)
    # (the oracle takes weights as complex64: the cross product is rounded)
    np.testing.assert_allclose(Em, Ee, rtol=1e-6, atol=1e-6 * np.abs(Ee).max())
    # electric rows are unchanged; magnetic line-of-sight fields are orthogonal
    # to the leg (the launch transform leaves w x d as it is)
    el = oracle.field_paths(pos, nrm, w, idx, S, K, 0.0, BETA)
    assert np.array_equal(Em[1::2], el[1::2])
    assert not np.allclose(Em[0::2], el[0::2])
