"""GPU tests of calculate_scattering(scattering=2): workchunkingv2 with
max_scatter 3 (rayfunctions.py:2863-2953) on the device raycaster, the
4-column path network of EMGPUFreqDomain (freqdomainkernal,
empropagation.py:1083-1265) and the leg-by-leg source sum, all against the
oracle's path list."""
import numpy as np
import pytest

from lyceanem.mesh import Mesh

pytestmark = pytest.mark.gpu


def _index4(rng, S, K, P, V):
    """Random valid 4-column rows: [s, k, 0, 0], [s, p, k, 0], [s, p, q, k]."""
    s = rng.integers(1, S + 1, V)
    k = rng.integers(S + 1, S + K + 1, V)
    p = rng.integers(S + K + 1, S + K + P + 1, V)
    q = rng.integers(S + K + 1, S + K + P + 1, V)
    q = np.where(q == p, S + K + 1 + (q - S - K) % P, q)
    kind = rng.integers(0, 3, V)
    rows = np.zeros((V, 4), np.int32)
    rows[:, 0] = s
    rows[kind == 0, 1] = k[kind == 0]
    rows[kind == 1, 1], rows[kind == 1, 2] = p[kind == 1], k[kind == 1]
    rows[kind == 2, 1], rows[kind == 2, 2], rows[kind == 2, 3] = p[kind == 2], q[kind == 2], \
        k[kind == 2]
    return rows


def test_field_paths_four_columns_vs_oracle(native, oracle):
    """lem_field_paths on 4-column rows (the zero-column wrap to the last point
    included) against the oracle's float64 path sum per (source, sink) slot."""
    from lyceanem.base_types import scattering_t
    from lyceanem.electromagnetics import empropagation as EM

    rng = np.random.default_rng(41)
    S, K, P = 6, 7, 9
    N = S + K + P
    pos = rng.uniform(-5, 5, (N, 3)).astype(np.float32)
    nrm = rng.normal(size=(N, 3)).astype(np.float32)
    w = (rng.normal(size=(N, 3)) + 1j * rng.normal(size=(N, 3))).astype(np.complex64)
    idx = _index4(rng, S, K, P, 4000)
    table = np.zeros(N, dtype=scattering_t)
    table["px"], table["py"], table["pz"] = pos.T
    table["nx"], table["ny"], table["nz"] = nrm.T
    table["ex"], table["ey"], table["ez"] = w.T
    table["Electric"] = True
    lam = 0.3
    got = EM.EMGPUFreqDomain(S, K, idx, table, lam, 0.01, 2 * np.pi / lam)
    ref = oracle.field_paths(pos, nrm, w, idx, S, K, 0.01, 2 * np.pi / lam)
    err = np.linalg.norm(got - ref) / np.linalg.norm(ref)
    per = np.abs(got - ref).max() / np.abs(ref).max()
    print(f"4-column paths: normwise {err:.2e}, max/max {per:.2e}", flush=True)
    assert err < 1e-5 and per < 1e-5


def _scene(n_tris, seed=17, S=4, K=12, P=30):
    rng = np.random.default_rng(seed)
    src = (rng.uniform(-0.3, 0.3, (S, 3)) * [1, 1, 0]).astype(np.float32)
    src_n = np.tile(np.array([[0, 0, 1]], np.float32), (S, 1))
    # the clutter spreads with the triangle count (about the same density)
    ext = 6.0 * np.sqrt(n_tris / 300.0)
    ctr = rng.uniform(-ext, ext, (n_tris, 1, 3)) * [1, 1, 0.4 * 6 / ext] + [0, 0, 6]
    tris = (ctr + rng.uniform(-0.6, 0.6, (n_tris, 3, 3))).reshape(-1, 9).astype(np.float32)
    t = tris.reshape(-1, 3, 3).astype(np.float64)
    fn = np.cross(t[:, 1] - t[:, 0], t[:, 2] - t[:, 0])
    fn /= np.linalg.norm(fn, axis=1, keepdims=True)
    pick = rng.choice(n_tris, P, replace=False)
    scat = (t[pick].mean(1) + 1e-3 * fn[pick]).astype(np.float32)
    scat_n = fn[pick].astype(np.float32)
    snk = (rng.uniform(-ext, ext, (K, 3)) * [1, 1, 0] + [0, 0, 14]).astype(np.float32)
    snk_n = np.tile(np.array([[0, 0, -1]], np.float32), (K, 1))
    w = (rng.normal(size=(S, 3)) + 1j * rng.normal(size=(S, 3))).astype(np.complex64)
    return dict(src=src, src_n=src_n, tris=tris, scat=scat, scat_n=scat_n, snk=snk,
                snk_n=snk_n, w=w)


def _oracle_paths(oracle, sc, w, scat_w, beta, los=True):
    S, K, P = len(sc["src"]), len(sc["snk"]), len(sc["scat"])
    full, _ = oracle.workchunking(sc["src"], sc["snk"], sc["scat"], sc["tris"], 3,
                                  line_of_sight=los)
    pos = np.concatenate([sc["src"], sc["snk"], sc["scat"]])
    nrm = np.concatenate([sc["src_n"], sc["snk_n"], sc["scat_n"]])
    wt = np.concatenate([np.asarray(w, np.complex64), np.ones((K, 3), np.complex64),
                         np.asarray(scat_w, np.complex64)])
    return oracle.field_paths(pos, nrm, wt, full, S, K, 0.0, beta), full


def _rel(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


@pytest.mark.parametrize("n_tris", [300, 5000])
def test_two_bounce_fields_vs_oracle(native, oracle, n_tris):
    """two_bounce_fields (brute-force masks below 4096 triangles, grid legs
    above) against the oracle's max_scatter-3 path list summed over sources."""
    from lyceanem.models import frequency_domain as FD

    sc = _scene(n_tris)
    beta = 2 * np.pi / 0.2
    scat_w = np.full((len(sc["scat"]), 3), 0.5, np.complex64)
    for los in (True, False):
        print(f"two bounces: {n_tris} triangles, los={los}", flush=True)
        net, full = _oracle_paths(oracle, sc, sc["w"], scat_w, beta, los=los)
        depth = (full[:, 1:] != 0).sum(1)
        assert (depth == 3).sum() > 0 and (depth == 2).sum() > 0
        ref = net.sum(0)
        got = FD.two_bounce_fields(sc["src"], sc["w"], sc["snk"], sc["snk_n"], sc["scat"],
                                   sc["scat_n"], sc["tris"], scat_w=scat_w, beta=beta, los=los)
        err = _rel(got, ref)
        big = np.abs(ref) >= 1e-3 * np.abs(ref).max()
        per = (np.abs(got - ref)[big] / np.abs(ref[big])).max()
        ph = np.abs(np.angle(got[big] / ref[big])).max()
        print(f"  {len(full)} paths ({(depth == 3).sum()} two-bounce): normwise {err:.2e}, "
              f"per-sink {per:.2e}, phase {ph:.2e} rad", flush=True)
        assert err < 1e-4 and per < 1e-4 and ph < 1e-3


def test_workchunking_three_matches_oracle(native, oracle, monkeypatch):
    """workchunkingv2(max_scatter=3) on the device raycaster equals the
    oracle's rows in order, also when the waves are split into memory chunks."""
    from lyceanem.raycasting import rayfunctions as RF

    sc = _scene(300, S=5, K=6, P=12)
    exp, erays = oracle.workchunking(sc["src"], sc["snk"], sc["scat"], sc["tris"], 3)
    got, rays = RF.workchunkingv2(sc["src"], sc["snk"], sc["scat"], sc["tris"], 3)
    assert np.array_equal(got, exp) and rays == erays
    monkeypatch.setattr(RF, "free_memory_bytes", lambda: 60000)
    got2, _ = RF.workchunkingv2(sc["src"], sc["snk"], sc["scat"], sc["tris"], 3)
    assert np.array_equal(got2, exp)


def test_calculate_scattering_two_bounces_dropin(native, oracle):
    """calculate_scattering(scattering=2) on meshes, summed (leg by leg) and
    per element (the 4-column path network), against the oracle."""
    from lyceanem.models import frequency_domain as FD

    sc = _scene(300, S=4, K=10, P=20)
    S, K, P = len(sc["src"]), len(sc["snk"]), len(sc["scat"])
    aperture = Mesh(points=sc["src"].astype(np.float64), cells=[],
                    point_data={"Normals": sc["src_n"], "Area": np.full(S, 3.0)})
    sinks = Mesh(points=sc["snk"].astype(np.float64), cells=[],
                 point_data={"Normals": sc["snk_n"]})
    scat = Mesh(points=sc["scat"].astype(np.float64), cells=[],
                point_data={"Normals": sc["scat_n"]})
    tv = sc["tris"].reshape(-1, 3).astype(np.float64)
    solid = Mesh(points=tv, cells=[("triangle", np.arange(len(tv)).reshape(-1, 3))])
    lam = 0.2
    w = np.tile(np.array([[0.0, 1.0, 0.5]], np.complex64), (S, 1))
    net, _ = _oracle_paths(oracle, sc, w, np.ones((P, 3)), 2 * np.pi / lam)
    for elements in (False, True):
        Ex, Ey, Ez = FD.calculate_scattering(aperture, sinks, solid, np.array([[0.0, 1.0, 0.5]]),
                                             scatter_points=scat, wavelength=lam, scattering=2,
                                             elements=elements, beta=2 * np.pi / lam)
        got = np.stack([Ex, Ey, Ez], -1)
        ref = net if elements else net.sum(0)
        assert got.shape == ref.shape and got.dtype == np.complex128
        err = _rel(got, ref)
        print(f"calculate_scattering(scattering=2, elements={elements}): normwise {err:.2e}",
              flush=True)
        assert err < 1e-4
