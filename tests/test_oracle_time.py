"""CPU checks that pin the time-domain oracle (timedomainkernal,
empropagation.py:1345-1612) with facts independent of its code: Python's own
float floor division, closed-form single- and two-leg amplitudes built from
projectors in numpy float64, the reference's wrap of a zero index entry, and
placement invariants (delta excitation, energy, truncation, wake)."""
import numpy as np
import pytest

C = 299792458.0


def _proj(d):
    d = np.asarray(d, np.float64)
    d = d / np.linalg.norm(d)
    return np.eye(3) - np.outer(d, d)


def _loss_mag(p_a, p_b, n_b, alpha, beta):
    """|-(1/2pi) (-(alpha + j beta) - 1/r) e^{-(alpha + j beta) r} / r (d.n)|,
    float64 throughout (the reference's float32 roundings stay below 1e-6)."""
    d = np.asarray(p_b, np.float64) - np.asarray(p_a, np.float64)
    r = np.linalg.norm(d)
    cos = abs(np.dot(d / r, n_b))
    return abs(alpha + 1j * beta + 1 / r) * np.exp(-alpha * r) / r * cos / (2 * np.pi)


def test_py_floordiv_matches_python(oracle):
    rng = np.random.default_rng(0)
    vals = [(0.3, 0.1), (7.0, 2.0), (-7.5, 2.0), (1e-9, 1e-9), (3e-9, 1e-9), (2.9999999e-9, 1e-9),
            (0.0, 1e-9), (5.0, 0.5), (1e-3, 1e-9), (123456.789, 1e-3)]
    for a, b in vals:
        assert oracle.py_floordiv(a, b) == a // b, (a, b)
    for _ in range(20000):
        b = 10.0 ** rng.uniform(-12, 0)
        a = b * rng.integers(0, 10 ** 6) * (1 + rng.choice([0.0, 1e-16, -1e-16, 1e-9]))
        assert oracle.py_floordiv(a, b) == a // b, (a, b)


def _table(rng, n):
    pos = rng.uniform(-20, 20, (n, 3)).astype(np.float32)
    nrm = rng.normal(size=(n, 3))
    nrm = (nrm / np.linalg.norm(nrm, axis=1, keepdims=True)).astype(np.float32)
    w = (rng.normal(size=(n, 3)) + 1j * rng.normal(size=(n, 3)) * 1e-9).astype(np.complex64)
    return pos, nrm, w


@pytest.mark.parametrize("alpha", [0.0, 0.01])
def test_los_path_closed_form(oracle, alpha):
    rng = np.random.default_rng(1)
    pos, nrm, w = _table(rng, 6)
    beta = 2 * np.pi / 0.05
    index = np.array([[1, 4], [2, 5], [3, 6]], np.int32)
    amp, arr = oracle.time_paths(pos, nrm, w, index, alpha, beta)
    for v, (a, b) in enumerate(index - 1):
        e = _proj(pos[b].astype(np.float64) - pos[a]) @ w[a].real.astype(np.float64)
        e = e * w[b].real
        m = _loss_mag(pos[a], pos[b], nrm[b].astype(np.float64), alpha, beta)
        np.testing.assert_allclose(amp[v], e * m, rtol=1e-6, atol=1e-12 * np.abs(e * m).max())
        r = np.linalg.norm(pos[b].astype(np.float64) - pos[a])
        delay = float(w[a, 0].imag) + float(w[b, 0].imag)
        # the reference subtracts coordinates in float32 (ulp ~2e-6 m here)
        assert arr[v] == pytest.approx(r / C + delay, rel=1e-6)


def test_arrival_exact_on_representable_geometry(oracle):
    """Coordinates whose float32 differences are exact: arrival = r / c + delay
    to float64 rounding, with r = 3, 5, 13 (Pythagorean triples)."""
    pos = np.array([[0, 0, 0], [3, 4, 0], [0, 0, 0], [0, 5, 12], [1, 1, 1], [4, 5, 1]], np.float32)
    nrm = np.tile(np.array([0, 0, 1], np.float32), (6, 1))
    w = np.ones((6, 3), np.complex64)
    w[:, 0] += 1j * np.array([0, 0, 2e-9, 0, 0, 1e-9], np.float32)
    amp, arr = oracle.time_paths(pos, nrm, w, np.array([[1, 2], [3, 4], [5, 6]], np.int32),
                                 0.0, 1.0)
    delay = [0.0, float(np.float32(2e-9)), float(np.float32(1e-9))]
    np.testing.assert_allclose(arr, np.array([5.0, 13.0, 5.0]) / C + delay, rtol=1e-15)


def test_one_bounce_closed_form(oracle):
    rng = np.random.default_rng(2)
    pos, nrm, w = _table(rng, 5)
    beta = 2 * np.pi / 0.1
    index = np.array([[1, 3, 2], [1, 4, 2], [5, 3, 2]], np.int32)
    amp, arr = oracle.time_paths(pos, nrm, w, index, 0.0, beta)
    P = pos.astype(np.float64)
    for v, (s, p, k) in enumerate(index - 1):
        e = w[s].real.astype(np.float64)
        e = (_proj(P[p] - P[s]) @ e) * w[p].real
        e = _proj(nrm[p]) @ e
        e = (_proj(P[k] - P[p]) @ e) * w[k].real
        m = (_loss_mag(P[s], P[p], nrm[p].astype(np.float64), 0.0, beta)
             * _loss_mag(P[p], P[k], nrm[k].astype(np.float64), 0.0, beta))
        np.testing.assert_allclose(amp[v], e * m, rtol=1e-6, atol=1e-9 * np.abs(e * m).max())
        r = np.linalg.norm(P[p] - P[s]) + np.linalg.norm(P[k] - P[p])
        delay = float(w[s, 0].imag) + float(w[p, 0].imag) + float(w[k, 0].imag)
        assert arr[v] == pytest.approx(r / C + delay, rel=1e-6)


def test_los_row_of_a_three_column_index_reads_the_last_point(oracle):
    """timedomainkernal reads point_information[0 - 1] for the zero column of a
    line-of-sight row: the leg length to the last point of the table is added
    and the field is projected onto the sink's tangent plane."""
    rng = np.random.default_rng(3)
    pos, nrm, w = _table(rng, 5)
    w.imag[:] = 0
    amp3, arr3 = oracle.time_paths(pos, nrm, w, np.array([[1, 2, 0]], np.int32), 0.0, 10.0)
    amp2, arr2 = oracle.time_paths(pos, nrm, w, np.array([[1, 2]], np.int32), 0.0, 10.0)
    P = pos.astype(np.float64)
    assert arr3[0] == pytest.approx(arr2[0] + np.linalg.norm(P[4] - P[1]) / C, rel=1e-6)
    np.testing.assert_allclose(amp3[0], _proj(nrm[1]) @ amp2[0], rtol=1e-9, atol=1e-15)


def test_placement_invariants(oracle):
    rng = np.random.default_rng(4)
    V, N, fs = 40, 400, 1e9
    amp = rng.normal(size=(V, 3))
    arrival = 1e-7 + rng.uniform(0, 2e-7, V)
    slot = rng.integers(0, 3, V)
    # delta excitation: one sample per path at its start, equal to amp
    out, wake, bins = oracle.time_place(amp, arrival, slot, 3, [1.0], fs, N)
    assert wake == arrival.min()
    assert bins.min() == 0
    np.testing.assert_array_equal(bins, [(a - wake) // (1 / fs) for a in arrival])
    ref = np.zeros((3, N, 3))
    for v in range(V):
        ref[slot[v], bins[v]] += amp[v]
    np.testing.assert_allclose(out, ref, rtol=1e-14, atol=1e-14)
    # a longer pulse: total per slot = sum(amp) * sum(excitation) while nothing is cut
    exc = rng.normal(size=50)
    out, _, bins = oracle.time_place(amp, arrival, slot, 3, exc, fs, N)
    assert bins.max() + 50 <= N
    for s in range(3):
        np.testing.assert_allclose(out[s].sum(0), amp[slot == s].sum(0) * exc.sum(), rtol=1e-12)
    # truncation at N: only the samples before N are placed
    out, _, bins = oracle.time_place(amp, arrival, slot, 3, exc, fs, 210)
    v = int(np.argmax(bins))
    kept = 210 - bins[v]
    assert 0 < kept < 50
    solo, _, _ = oracle.time_place(amp[v:v + 1], arrival[v:v + 1], [0], 1, exc, fs, 210 + 0)
    # alone, the path is the earliest arrival, so it starts at sample 0
    np.testing.assert_allclose(solo[0, :50], np.outer(exc, amp[v]), rtol=1e-14)


def test_wake_starts_at_one_second(oracle):
    """The reference's wake time starts at 1 s: arrivals later than that are
    placed relative to 1 s."""
    amp = np.ones((2, 3))
    arrival = np.array([2.0, 2.0 + 5e-9])
    out, wake, bins = oracle.time_place(amp, arrival, [0, 0], 1, [1.0], 1e9, 10)
    assert wake == 1.0
    assert (bins > 10).all() and not out.any()


def test_magnetic_first_point(oracle):
    """A magnetic first point launches Re(w) x d / eta (empropagation.py:
    1395-1419): equal to an electric point of weight numpy.cross(Re w, d) / eta
    with the same delay; w = z towards +x gives +y."""
    eta0 = float(np.sqrt(1.25663706212e-6 / 8.8541878187e-12))
    pos = np.array([[0, 0, 0], [200, 0, 0]], np.float32)
    nrm = np.array([[0, 0, 1], [1, 0, 0]], np.float32)
    w = np.array([[0, 0, 1], [1, 1, 1]], np.complex64)
    amp, _ = oracle.time_paths(pos, nrm, w, np.array([[1, 2]], np.int32), 0.0, 2 * np.pi / 0.05,
                               eta=np.array([eta0, 0.0]))
    assert amp[0, 1] > 0 and amp[0, 0] == 0 and amp[0, 2] == 0
    rng = np.random.default_rng(5)
    pos, nrm, w = _table(rng, 12)
    eta = np.where(np.arange(12) % 3 == 0, eta0, 0.0)
    idx = np.array([[a, b, c] for a in range(1, 5) for b in range(5, 9) for c in (0, 10)],
                   np.int32)
    beta = 2 * np.pi / 0.05
    am, arm = oracle.time_paths(pos, nrm, w, idx, 0.01, beta, eta=eta)
    for v, row in enumerate(idx):
        a, b = row[0] - 1, row[1] - 1
        we = w.copy()
        if eta[a] > 0:
            d = pos[b].astype(np.float64) - pos[a]
            d /= np.linalg.norm(d)
            we[a] = (np.cross(w[a].real.astype(np.float64), d) / eta[a]) + 1j * w[a].imag
        ae, are = oracle.time_paths(pos, nrm, we, row[None, :], 0.01, beta)
        np.testing.assert_allclose(am[v], ae[0], rtol=1e-6, atol=1e-9 * np.abs(ae).max())
        assert arm[v] == are[0]
