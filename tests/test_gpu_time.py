"""GPU tests of the time-domain path (lem_time.hip) against the oracle:
per-path amplitudes and arrival times, the wake time, start samples (Python's
float floor division), the per-slot time map and the summed trace, and the
drop-ins TimeDomainv3 / models.time_domain.calculate_scattering."""
import numpy as np
import pytest

from lyceanem.mesh import Mesh

pytestmark = pytest.mark.gpu


def _table(rng, n, delays=True):
    pos = rng.uniform(-20, 20, (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    im = rng.normal(size=(n, 3)) * (1e-9 if delays else 0.0)
    w = (rng.normal(size=(n, 3)) + 1j * im).astype(np.complex64)
    return pos, nrm, w


def _index(rng, n, V, ncol):
    idx = rng.integers(1, n + 1, (V, ncol)).astype(np.int32)
    for c in range(1, ncol):
        # no zero-length legs (the reference's loss is then infinite too)
        same = idx[:, c] == idx[:, c - 1]
        idx[same, c] = idx[same, c] % n + 1
    if ncol > 2:
        # trailing zero columns on part of the rows (line-of-sight rows of a
        # wider index), which the reference reads as point -1
        cut = rng.integers(2, ncol + 1, V)
        for c in range(2, ncol):
            idx[cut <= c, c] = 0
    return idx


@pytest.mark.parametrize("ncol,magnetic", [(2, False), (3, False), (4, False), (3, True)])
def test_time_paths_match_oracle(native, oracle, ncol, magnetic):
    """Per-path amplitudes and arrivals; with magnetic=True every third point
    is magnetic (impedance 376.73 ohm), launching Re(w) x d / eta."""
    print(f"time paths ncol={ncol} magnetic={magnetic}", flush=True)
    rng = np.random.default_rng(ncol)
    pos, nrm, w = _table(rng, 50)
    idx = _index(rng, 50, 3000, ncol)
    eta = np.where(np.arange(50) % 3 == 0, np.float32(376.73), 0.0) if magnetic else None
    amp, arr = native.time_paths(pos, nrm, w, idx, 0.02, 2 * np.pi / 0.03, eta=eta)
    amp_o, arr_o = oracle.time_paths(pos, nrm, w, idx, 0.02, 2 * np.pi / 0.03, eta=eta)
    amp, arr = amp.cpu().numpy(), arr.cpu().numpy()
    assert np.array_equal(np.isnan(amp), np.isnan(amp_o))
    scale = np.nanmax(np.abs(amp_o))
    np.testing.assert_allclose(amp, amp_o, rtol=1e-10, atol=1e-13 * scale)
    np.testing.assert_allclose(arr, arr_o, rtol=1e-14, atol=0)


def test_time_paths_reject_bad_index(native):
    rng = np.random.default_rng(0)
    pos, nrm, w = _table(rng, 5)
    with pytest.raises(native.LemError, match="outside"):
        native.time_paths(pos, nrm, w, np.array([[1, 6]], np.int32), 0.0, 1.0)
    with pytest.raises(native.LemError, match="outside"):
        native.time_paths(pos, nrm, w, np.array([[0, 2]], np.int32), 0.0, 1.0)


def test_wake_and_bins_exact(native):
    rng = np.random.default_rng(5)
    arrival = 3e-7 + rng.uniform(0, 1e-6, 5000)
    arrival[17] = 1e-7
    wake = native.time_wake(arrival).item()
    assert wake == 1e-7
    fs = 2.5e9
    bins = native.time_bins(arrival, np.array([wake]), fs).cpu().numpy()
    exp = np.array([(a - wake) // (1 / fs) for a in arrival], np.int64)
    np.testing.assert_array_equal(bins, exp)
    # negative arrivals (negative delays) and the 1 s cap
    assert native.time_wake(np.array([-2e-9, 5e-9])).item() == -2e-9
    assert native.time_wake(np.array([3.0, 4.0])).item() == 1.0


@pytest.mark.parametrize("E,N", [(1, 300), (64, 500), (64, 120)])
def test_time_map_and_trace_match_oracle(native, oracle, E, N):
    print(f"time map E={E} N={N}", flush=True)
    rng = np.random.default_rng(E + N)
    V, nslots = 4000, 37
    amp = rng.normal(size=(V, 3))
    arrival = 1e-7 + rng.uniform(0, 3e-7, V)
    slot = rng.integers(0, nslots, V)
    exc = rng.normal(size=E)
    fs = 1e9
    ref, wake_o, bins_o = oracle.time_place(amp, arrival, slot, nslots, exc, fs, N)
    wake = native.time_wake(arrival)
    assert wake.item() == wake_o
    bins = native.time_bins(arrival, wake, fs)
    np.testing.assert_array_equal(bins.cpu().numpy(), bins_o)
    tm = native.time_map(amp, bins, slot, nslots, exc, N).cpu().numpy()
    np.testing.assert_allclose(tm, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max())
    tr = native.time_trace(amp, bins, exc, N).cpu().numpy()
    np.testing.assert_allclose(tr, ref.sum(0), rtol=1e-11, atol=1e-11 * np.abs(ref).max())


def _scene(rng):
    # 4 x 4 array at z = 0 facing +z, 6 sinks 30-50 m up, a few blocking plates
    g = np.arange(4) * 0.05
    X, Y = np.meshgrid(g, g)
    src = np.stack([X.ravel(), Y.ravel(), np.zeros(16)], 1)
    snk = np.stack([rng.uniform(-20, 20, 6), rng.uniform(-20, 20, 6), rng.uniform(30, 50, 6)], 1)
    tris = []
    for _ in range(12):
        c = np.array([rng.uniform(-8, 8), rng.uniform(-8, 8), rng.uniform(5, 20)])
        tris.append(c + rng.uniform(-3, 3, (3, 3)))
    tris = np.array(tris)
    blocker = Mesh(points=tris.reshape(-1, 3), cells=[("triangle", np.arange(36).reshape(12, 3))])
    ap = Mesh(points=src, point_data={"Normals": np.tile([0.0, 0, 1], (16, 1)),
                                      "Area": np.full(16, 0.0025)})
    sk = Mesh(points=snk, point_data={"Normals": -snk / np.linalg.norm(snk, axis=1,
                                                                        keepdims=True)})
    return ap, sk, blocker, tris.reshape(-1, 9).astype(np.float32)


def _oracle_model(oracle, ap, sk, tris9, w, exc, fs, N, alpha, beta):
    src = np.asarray(ap.points, np.float32)
    snk = np.asarray(sk.points, np.float32)
    S, K = len(src), len(snk)
    fi, _ = oracle.workchunking(src, snk, np.empty((0, 3), np.float32), tris9, 1)
    pos = np.concatenate([src, snk]).astype(np.float32)
    nrm = np.concatenate([ap.point_data["Normals"], sk.point_data["Normals"]]).astype(np.float32)
    ww = np.concatenate([w, np.ones((K, 3))]).astype(np.complex64)
    amp, arr = oracle.time_paths(pos, nrm, ww, fi, alpha, beta)
    s = fi[:, 0] - 1
    return fi, amp, arr, s


def test_time_domain_model_matches_oracle(native, oracle):
    from lyceanem.models import time_domain

    rng = np.random.default_rng(11)
    ap, sk, blocker, tris9 = _scene(rng)
    fs, N = 4e9, 900
    exc = np.sin(np.linspace(0, 6 * np.pi, 40)) * np.hanning(40)
    beta = 2 * np.pi / 0.1
    Ex, Ey, Ez, wake = time_domain.calculate_scattering(
        ap, sk, exc, blocker, np.array([[1.0, 0, 0]]), wavelength=0.1, sampling_freq=fs,
        num_samples=N, project_vectors=False, beta=beta)
    w = np.repeat(np.array([[1.0, 0, 0]], np.complex64), 16, 0) * 0.0025
    fi, amp, arr, s = _oracle_model(oracle, ap, sk, tris9, w, exc, fs, N, 0.0, beta)
    assert 0 < len(fi) < 16 * 6  # some rays blocked, some not
    ref, wake_o, _ = oracle.time_place(amp, arr, np.zeros(len(fi), np.int64), 1, exc, fs, N)
    assert wake == wake_o
    wi = int(np.digitize(wake_o, np.linspace(0, N / fs, N)))
    tr = np.zeros((N, 3))
    tr[wi:] = ref[0, :N - wi]
    for out, c in ((Ex, 0), (Ey, 1), (Ez, 2)):
        assert out.shape == (N,)
        np.testing.assert_allclose(out, tr[:, c], rtol=1e-9, atol=1e-12 * np.abs(tr).max())
    assert np.abs(Ex).max() > 0

    # elements=True: every sink row of element e is the trace of e alone / S
    Exe, _, _, _ = time_domain.calculate_scattering(
        ap, sk, exc, blocker, np.array([[1.0, 0, 0]]), wavelength=0.1, sampling_freq=fs,
        num_samples=N, project_vectors=False, beta=beta, elements=True)
    assert Exe.shape == (16, 6, N)
    per, _, _ = oracle.time_place(amp / 16, arr, s, 16, exc, fs, N)
    for e in (0, 7, 15):
        for k in (0, 5):
            np.testing.assert_allclose(Exe[e, k], per[e, :, 0], rtol=1e-9,
                                       atol=1e-12 * np.abs(per).max())


def test_timedomainv3_dropin(native, oracle):
    from lyceanem.electromagnetics import empropagation as EM
    from lyceanem.models.frequency_domain import _point_table

    rng = np.random.default_rng(12)
    pos, nrm, w = _table(rng, 12)
    S, K = 4, 5
    # rows source -> sink, and source -> scatterer -> sink
    rows = [[s + 1, S + k + 1, 0] for s in range(S) for k in range(K) if (s + k) % 3]
    rows += [[s + 1, S + K + p + 1, S + k + 1] for s in range(S) for p in range(3)
             for k in range(K) if (s * p + k) % 4 == 0]
    fi = np.array(rows, np.int32)
    table = _point_table([pos], [nrm], [w])
    exc = rng.normal(size=30)
    tm, wake = EM.TimeDomainv3(S, K, table, fi, 0.25, 0.1, exc, 2e9, 400, 0.0, 20.0)
    amp, arr = oracle.time_paths(pos, nrm, w, fi, 0.0, 20.0)
    s, k = EM.path_slots(fi, S, K)
    ref, wake_o, _ = oracle.time_place(amp, arr, s * K + k, S * K, exc, 2e9, 400)
    assert tm.shape == (S, K, 400, 3) and wake == wake_o
    np.testing.assert_allclose(tm, ref.reshape(S, K, 400, 3), rtol=1e-9,
                               atol=1e-12 * np.abs(ref).max())
    # magnetic sources 0 and 2 (free-space impedance of the table's constants)
    table["Electric"][[0, 2]] = False
    eta = np.zeros(12)
    eta[[0, 2]] = np.float32(np.sqrt(np.float64(table["permeability"][0].real)
                                     / np.float64(table["permittivity"][0].real)))
    tm, wake = EM.TimeDomainv3(S, K, table, fi, 0.25, 0.1, exc, 2e9, 400, 0.0, 20.0)
    amp, arr = oracle.time_paths(pos, nrm, w, fi, 0.0, 20.0, eta=eta)
    ref, wake_o, _ = oracle.time_place(amp, arr, s * K + k, S * K, exc, 2e9, 400)
    assert wake == wake_o and np.abs(ref).max() > 0
    np.testing.assert_allclose(tm, ref.reshape(S, K, 400, 3), rtol=1e-9,
                               atol=1e-12 * np.abs(ref).max())
