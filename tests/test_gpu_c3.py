"""GPU tests of the one-bounce scattering network (the 256 x 256 array, 50k
triangle, 200k scatterer, 64 x 64 sink workload) on the path it is benchmarked
with: visibility through the uniform grid on every leg (leg 1 sink-major and
walked from the scatterer end, leg 2 source-major, line of sight sink-major)
followed by the masked field kernel.

The reference computes the same sum as a path list: workchunkingv2 with
max_scatter 2 (rayfunctions.py:2767-2861) and freqdomainkernal summed over the
sources (frequency_domain.py:596-649).  The oracle restates that pipeline on the
CPU (brute-force kernel1Dv3 masks, float64 fields), so a reduced scene with
enough triangles to take the grid path is compared with it end to end, and the
full-size scene's masks are compared ray by ray on a random sample."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _reduced_scene(n_src_side=4, n_scat=700, n_snk_side=10):
    from lyceanem import scenes

    # 5,000 triangles (above the grid threshold, so every leg takes the grid
    # path); fewer sources, scatterers and sinks than the full workload, the
    # same geometry otherwise
    return scenes.clutter_c3(n_src_side=n_src_side, n_tris=5000, n_scat=n_scat,
                             n_snk_side=n_snk_side)


def _oracle_one_bounce(oracle, sc, w, scat_w, los=True):
    """Sum over sources of the oracle's path network: workchunkingv2(..., 2)
    rows (every second-leg ray cast per source, as the reference does), then
    freqdomainkernal per path (float64)."""
    src, snk, scat = sc["src"], sc["snk"], sc["scat"]
    S, K, P = len(src), len(snk), len(scat)
    full, _ = oracle.workchunking(src, snk, scat, sc["tris"], 2, line_of_sight=los)
    pos = np.concatenate([src, snk, scat]).astype(np.float32)
    nrm = np.concatenate([sc["src_n"], sc["snk_n"], sc["scat_n"]]).astype(np.float32)
    wt = np.concatenate([np.asarray(w, np.complex64).reshape(S, 3),
                         np.ones((K, 3), np.complex64),
                         np.asarray(scat_w, np.complex64).reshape(P, 3)])
    net = oracle.field_paths(pos, nrm, wt, full, S, K, sc["alpha"], sc["beta"])
    return net.sum(0), full


def _oracle_legs(oracle, sc, w, scat_w, los=True):
    """The same sum leg by leg (brute-force masks, float64 fields): the
    scatterer fields H_p, then H -> sinks plus the line of sight.  Equal to the
    path sum by linearity (pinned on the CPU against field_paths); H enters the
    second leg rounded to complex64."""
    src, snk, scat, tris = sc["src"], sc["snk"], sc["scat"], sc["tris"]
    a, b = sc["alpha"], sc["beta"]
    h1, _ = oracle.visibility_dense(src, scat, tris, any_hit_only=True, want_dist=False)
    assert (~h1).any()
    H = oracle.field_los(src, w, scat, sc["scat_n"], a, b, hit=h1, snk_w=scat_w,
                         sink_tangent=True)
    h2, _ = oracle.visibility_dense(scat, snk, tris, any_hit_only=True, want_dist=False)
    E = oracle.field_los(scat, H, snk, sc["snk_n"], a, b, hit=h2)
    if los:
        h3, _ = oracle.visibility_dense(src, snk, tris, any_hit_only=True, want_dist=False)
        E = E + oracle.field_los(src, w, snk, sc["snk_n"], a, b, hit=h3, sink_tangent=True)
    return E, H, (h1.mean(), h2.mean())


def _errors(got, ref):
    """(normwise relative error, max per-sink relative error over sinks with
    |E_k| >= 1e-3 max |E|, max phase error in rad over the same sinks)."""
    got = np.asarray(got).reshape(-1)
    ref = np.asarray(ref).reshape(-1)
    norm = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    big = np.abs(ref) >= 1e-3 * np.abs(ref).max()
    per = np.abs(got[big] - ref[big]) / np.abs(ref[big])
    ph = np.abs(np.angle(got[big] / ref[big]))
    return float(norm), float(per.max()), float(ph.max())


def test_c3_reduced_grid_legs_vs_oracle_paths(native, oracle):
    """one_bounce_fields on the grid legs (scene weights, reflection 0.5) equals
    the oracle's workchunkingv2 + freqdomainkernal path sum within 1e-4."""
    from lyceanem.models import frequency_domain as FD

    sc = _reduced_scene()
    assert len(sc["tris"]) >= FD.GRID_MIN_TRIANGLES  # the grid legs run
    for los in (True, False):
        print(f"c3 reduced (los={los}): oracle path list", flush=True)
        ref, full = _oracle_one_bounce(oracle, sc, sc["w"], sc["scat_w"], los=los)
        n_scat_rows = int((full[:, 2] != 0).sum())
        assert n_scat_rows > 0
        got = FD.one_bounce_fields(sc["src"], sc["w"], sc["snk"], sc["snk_n"], sc["scat"],
                                   sc["scat_n"], sc["tris"], scat_w=sc["scat_w"],
                                   alpha=sc["alpha"], beta=sc["beta"], los=los)
        norm, per, ph = _errors(got, ref)
        print(f"c3 reduced (los={los}): {len(full)} paths ({n_scat_rows} via a scatterer); "
              f"normwise {norm:.2e}, per-sink max {per:.2e}, phase max {ph:.2e} rad", flush=True)
        assert norm < 1e-4 and per < 1e-4 and ph < 1e-3


def test_c3_grid_legs_vs_oracle_legs(native, oracle):
    """32 x 32 sources, 5,000 triangles, 2,000 scatterers, 16 x 16 sinks: the
    grid legs against the oracle leg by leg (brute-force masks, float64), at
    the scatterers (H) and at the sinks."""
    from lyceanem.models import frequency_domain as FD

    sc = _reduced_scene(n_src_side=32, n_scat=2000, n_snk_side=16)
    print("c3 32x32: oracle legs", flush=True)
    ref, H_ref, blocked = _oracle_legs(oracle, sc, sc["w"], sc["scat_w"])
    print("c3 32x32: grid legs", flush=True)
    got = FD.one_bounce_fields(sc["src"], sc["w"], sc["snk"], sc["snk_n"], sc["scat"],
                               sc["scat_n"], sc["tris"], scat_w=sc["scat_w"], alpha=sc["alpha"],
                               beta=sc["beta"])
    norm, per, ph = _errors(got, ref)
    # the scatterer fields of leg 1 alone, as the benchmark casts them
    g = native.Grid(sc["tris"])
    H = native.field_los_grid_sink_major(g, sc["src"], sc["w"], sc["scat"], sc["scat_n"],
                                         snk_w=sc["scat_w"], alpha=sc["alpha"], beta=sc["beta"],
                                         sink_tangent=True).cpu().numpy()
    hn, hp, hph = _errors(H, H_ref)
    print(f"c3 32x32: blocked leg 1 {blocked[0]:.3f}, leg 2 {blocked[1]:.3f}; sinks: normwise "
          f"{norm:.2e}, per-sink {per:.2e}, phase {ph:.2e} rad; scatterers: normwise {hn:.2e}, "
          f"per-scatterer {hp:.2e}, phase {hph:.2e} rad", flush=True)
    assert 0.05 < blocked[0] < 0.95 and 0.05 < blocked[1] < 0.95
    assert norm < 1e-4 and per < 1e-4 and ph < 1e-3
    assert hn < 1e-4 and hp < 1e-4 and hph < 1e-3


def test_c3_reduced_calculate_scattering_vs_oracle(native, oracle):
    """calculate_scattering(scattering=1) on meshes (the drop-in entry point):
    the reference's weights (real part of row 0 of desired_E_axis, no Area,
    scatterer weights 1; frequency_domain.py:452-485, 513-515) through the grid
    legs, against the oracle's path network."""
    from lyceanem.mesh import Mesh
    from lyceanem.models import frequency_domain as FD

    sc = _reduced_scene()
    S, K, P = len(sc["src"]), len(sc["snk"]), len(sc["scat"])
    aperture = Mesh(points=sc["src"].astype(np.float64), cells=[],
                    point_data={"Normals": sc["src_n"], "Area": np.full(S, 7.0)})
    sinks = Mesh(points=sc["snk"].astype(np.float64), cells=[],
                 point_data={"Normals": sc["snk_n"]})
    scat = Mesh(points=sc["scat"].astype(np.float64), cells=[],
                point_data={"Normals": sc["scat_n"]})
    tv = sc["tris"].reshape(-1, 3).astype(np.float64)
    solid = Mesh(points=tv, cells=[("triangle", np.arange(len(tv)).reshape(-1, 3))])
    desired = np.array([[0.6 + 0.3j, 0.8, 0.0], [5.0, 5.0, 5.0]])
    lam = sc["wavelength"]
    Ex, Ey, Ez = FD.calculate_scattering(aperture, sinks, solid, desired, scatter_points=scat,
                                         wavelength=lam, scattering=1, beta=2 * np.pi / lam)
    w = np.tile(np.array([[0.6, 0.8, 0.0]], np.complex64), (S, 1))
    ref, _ = _oracle_one_bounce(oracle, sc, w, np.ones((P, 3), np.complex64))
    got = np.stack([Ex, Ey, Ez], -1)
    assert got.dtype == np.complex128 and got.shape == (K, 3)
    norm, per, ph = _errors(got, ref)
    print(f"c3 reduced calculate_scattering: normwise {norm:.2e}, per-sink {per:.2e}, "
          f"phase {ph:.2e} rad", flush=True)
    assert norm < 1e-4 and per < 1e-4 and ph < 1e-3


def _bit(bits_np, rows, cols):
    """Bits (rows[i], cols[i]) of a row-major uint64 word array."""
    b = bits_np.view(np.uint64)
    words = b[rows, cols >> 6]
    return ((words >> (cols & 63).astype(np.uint64)) & np.uint64(1)).astype(bool)


def test_c3_full_scene_sampled_masks(native, oracle):
    """Full-size scene (65,536 sources, 50k triangles, 200k scatterers, 4,096
    sinks): the three grid masks exactly as the benchmark casts them, checked
    on 100,000 random leg-1 rays, 100,000 leg-2 rays and 20,000 line-of-sight
    rays against the oracle's brute force over every triangle."""
    from lyceanem import scenes

    sc = scenes.clutter_c3()
    src, scat, snk, tris = sc["src"], sc["scat"], sc["snk"], sc["tris"]
    S, P, K = len(src), len(scat), len(snk)
    g = native.Grid(tris)
    rng = np.random.default_rng(33)
    print("c3 full: leg-1 mask (sink-major, from the scatterer end)", flush=True)
    b1, _ = g.visibility(src, scat, any_hit=True, sink_major=True, from_sink=True)
    b1 = b1.cpu().numpy()
    n = 100_000
    s_i, p_i = rng.integers(0, S, n), rng.integers(0, P, n)
    got1 = _bit(b1, p_i, s_i)
    del b1
    print("c3 full: leg-2 mask (source-major) and line-of-sight mask (sink-major)", flush=True)
    b2, _ = g.visibility(scat, snk, any_hit=True)
    b2 = b2.cpu().numpy()
    q_i, k_i = rng.integers(0, P, n), rng.integers(0, K, n)
    got2 = _bit(b2, q_i, k_i)
    b3, _ = g.visibility(src, snk, any_hit=True, sink_major=True)
    b3 = b3.cpu().numpy()
    m = 20_000
    s3, k3 = rng.integers(0, S, m), rng.integers(0, K, m)
    got3 = _bit(b3, k3, s3)
    print("c3 full: oracle brute force on the samples", flush=True)
    ref1, _ = oracle.visibility_sampled(src, scat, np.stack([s_i, p_i], 1), tris,
                                        any_hit_only=True)
    ref2, _ = oracle.visibility_sampled(scat, snk, np.stack([q_i, k_i], 1), tris,
                                        any_hit_only=True)
    ref3, _ = oracle.visibility_sampled(src, snk, np.stack([s3, k3], 1), tris,
                                        any_hit_only=True)
    mm = [int((got1 != ref1).sum()), int((got2 != ref2).sum()), int((got3 != ref3).sum())]
    print(f"c3 full: mismatches leg 1 {mm[0]} / {n} (blocked {ref1.mean():.3f}), leg 2 "
          f"{mm[1]} / {n} (blocked {ref2.mean():.3f}), line of sight {mm[2]} / {m} "
          f"(blocked {ref3.mean():.3f})", flush=True)
    assert 0.05 < ref1.mean() < 0.95 and 0.05 < ref2.mean() < 0.95
    assert mm == [0, 0, 0]
