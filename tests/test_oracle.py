"""Pins for the FP64 CPU oracle (oracle/ssa_oracle.c).

Each check is chosen so that a plausible slip in the oracle (a dropped term,
a wrong sign, a transposed operand, an off-by-one extraction) fails it:
  * Philox4x32-10 against the Random123 known-answer vectors;
  * the FFT against numpy.fft and an O(n^2) DFT;
  * matvecs against brute-force double sums (MKL/test/ssa_opt_test.c:1464-1559);
  * the randomized SVD against a dense LAPACK SVD of the explicit Hankel
    matrix, exactly for a finite-rank signal and at cfg1 with q = 2;
  * reconstruction against explicit anti-diagonal averaging and the
    full-rank identity sum_i X_i = x;
  * W-correlation against its definition on explicit reconstructions;
  * the LRF against the sinusoid recurrence x[n] = 2 cos(w) x[n-1] - x[n-2];
  * the V-forecast against the same sinusoid and the LRF recurrence;
  * the reference's own ssa_opt_ref.h, compiled from source (oracle/_ref).
"""
import ctypes

import numpy as np
import pytest

import oracle as O
import workloads


def hankel(x, L):
    K = len(x) - L + 1
    return np.lib.stride_tricks.sliding_window_view(x, K)[:L].copy()


def diag_average(Xm):
    L, K = Xm.shape
    N = L + K - 1
    out = np.zeros(N)
    cnt = np.zeros(N)
    for i in range(L):
        out[i:i + K] += Xm[i]
        cnt[i:i + K] += 1
    return out / cnt


def test_philox_kat():
    # Random123 kat_vectors, philox4x32_10
    assert O.philox4x32_10([0, 0, 0, 0], [0, 0]) == [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]
    assert O.philox4x32_10([0xffffffff] * 4, [0xffffffff] * 2) == [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]
    assert O.philox4x32_10([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344], [0xa4093822, 0x299f31d0]) == \
        [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]


def test_philox_gaussian_moments():
    g = O.philox_gaussian(7, 3, 4001, 25)
    assert g.shape == (4001, 25)
    assert abs(g.mean()) < 0.01 and abs(g.std() - 1.0) < 0.01
    # different call index -> different matrix; same inputs -> identical
    assert not np.array_equal(g, O.philox_gaussian(7, 4, 4001, 25))
    assert np.array_equal(g, O.philox_gaussian(7, 3, 4001, 25))


@pytest.mark.parametrize("n", [2, 8, 64, 1024, 16384])
def test_fft_vs_numpy(n):
    rng = np.random.default_rng(n)
    z = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    ref = np.fft.fft(z)
    assert np.max(np.abs(O.fft(z, -1) - ref)) <= 1e-14 * np.max(np.abs(ref)) * np.log2(n + 1)
    assert np.allclose(O.fft(z, +1), np.conj(np.fft.fft(np.conj(z))), rtol=0, atol=1e-12 * np.abs(ref).max())


def test_fft_dft_bruteforce():
    n = 32
    z = np.random.default_rng(0).standard_normal(n) + 0j
    k = np.arange(n)
    D = np.exp(-2j * np.pi * np.outer(k, k) / n)
    assert np.max(np.abs(O.fft(z) - D @ z)) < 1e-13


@pytest.mark.parametrize("N,L", [(100, 25), (100, 75), (37, 2), (37, 36), (1000, 250)])
def test_matvec_bruteforce(N, L):
    x = np.random.default_rng(N + L).standard_normal(N)
    s = O.OracleSSA(x, L)
    H = hankel(x, L)
    v = np.random.default_rng(1).standard_normal(s.K)
    u = np.random.default_rng(2).standard_normal(s.L)
    scale = np.abs(x).max() * max(L, s.K)
    assert np.max(np.abs(s.matvec(v) - H @ v)) < 1e-13 * scale
    assert np.max(np.abs(s.matvec_T(u) - H.T @ u)) < 1e-13 * scale
    V = np.random.default_rng(3).standard_normal((s.K, 5))
    assert np.max(np.abs(s.matvec_block(V) - H @ V)) < 1e-13 * scale
    Ub = np.random.default_rng(4).standard_normal((s.L, 3))
    assert np.max(np.abs(s.matvec_block(Ub, transpose=True) - H.T @ Ub)) < 1e-13 * scale


def test_qr_and_jacobi():
    A = np.random.default_rng(0).standard_normal((50, 7))
    A[:, 3] *= 1e-6
    Q = np.asfortranarray(A.copy())
    R = np.zeros((7, 7), order="F")
    O.lib().orc_qr_q(O._d(Q), 50, 7, O._d(R))
    assert np.abs(Q.T @ Q - np.eye(7)).max() < 1e-14
    assert np.abs(Q @ R - A).max() < 1e-14 * np.abs(A).max()
    B = np.asfortranarray(A.copy())
    S = np.zeros(7)
    W = np.zeros((7, 7), order="F")
    O.lib().orc_jacobi_svd(O._d(B), 50, 7, O._d(S), O._d(W))
    sv = np.linalg.svd(A, compute_uv=False)
    # LAPACK's bidiagonal SVD is accurate to eps*sigma_1 absolutely
    assert np.allclose(S, sv, rtol=1e-13, atol=1e-14 * sv[0])
    assert np.abs((B * S) @ W.T - A).max() < 1e-14 * np.abs(A).max()


def finite_rank_signal(N):
    t = np.arange(N, dtype=np.float64)
    return 3.0 + 0.02 * t + 2 * np.sin(2 * np.pi * t / 17) + 0.7 * np.cos(2 * np.pi * t / 40 + 1) \
        + 0.3 * np.exp(-t / 300.0)


def test_rsvd_exact_on_finite_rank_signal():
    # rank 2 (linear) + 2 + 2 + 1 = 7 <= b: the range finder is exact for any Omega
    N, L, k, p = 400, 100, 7, 5
    x = finite_rank_signal(N)
    s = O.OracleSSA(x, L).decompose(k, p, 0, O.philox_gaussian(9, 0, N - L + 1, k + p))
    sv = np.linalg.svd(hankel(x, L), compute_uv=False)
    assert np.allclose(s.sigma, sv[:k], rtol=1e-11, atol=0)
    assert np.max(np.abs(s.reconstruct(list(range(k))) - x)) < 1e-11 * np.abs(x).max()
    U, V, sig = s.factors()
    assert np.abs(U.T @ U - np.eye(k)).max() < 1e-12
    assert np.abs(V.T @ V - np.eye(k)).max() < 1e-12
    # sign convention of MKL/ssa_opt.h:1749-1759
    assert np.all(U.sum(axis=0) >= 0)


def test_rsvd_cfg1_dense_oracle():
    x = workloads.single(1000, seed=1)
    L, k, p = 250, 50, 10
    K = len(x) - L + 1
    Uf, sv, Vt = np.linalg.svd(hankel(x, L), full_matrices=False)
    s = O.OracleSSA(x, L).decompose(k, p, 2, O.philox_gaussian(11, 0, K, k + p))
    rel = np.abs(s.sigma - sv[:k]) / sv[:k]
    # the signal components are resolved to rounding; the noise tail is a
    # randomized approximation (q = 2), bounded but not exact
    assert rel[:5].max() < 1e-10
    assert rel.max() < 0.2 and np.all(s.sigma <= sv[:k] * (1 + 1e-12))
    # near-degenerate pair [1,2] compared as a group (rotation invariant).
    # With q = 2 the subspace error (sigma_61/sigma_2)^5 ~ 1e-9 is the
    # algorithm's, not rounding; with q = 8 it vanishes below rounding.
    Xd = diag_average((Uf[:, 1:3] * sv[1:3]) @ Vt[1:3])
    Xd0 = diag_average(np.outer(Uf[:, 0] * sv[0], Vt[0]))
    rms = np.sqrt(np.mean(x ** 2))
    assert np.abs(s.reconstruct([1, 2]) - Xd).max() / rms < 1e-6
    s8 = O.OracleSSA(x, L).decompose(k, p, 8, O.philox_gaussian(11, 0, K, k + p))
    assert np.abs(s8.reconstruct([1, 2]) - Xd).max() / rms < 1e-10
    assert np.abs(s8.reconstruct([0]) - Xd0).max() / rms < 1e-10
    assert (np.abs(s8.sigma[:6] - sv[:6]) / sv[:6]).max() < 1e-12


def test_rsvd_converges_with_q():
    x = workloads.single(1000, seed=1)
    L, k, p = 250, 20, 10
    K = len(x) - L + 1
    sv = np.linalg.svd(hankel(x, L), compute_uv=False)
    errs = []
    for q in (0, 1, 2, 4):
        s = O.OracleSSA(x, L).decompose(k, p, q, O.philox_gaussian(11, 0, K, k + p))
        errs.append(np.max(np.abs(s.sigma - sv[:k]) / sv[:k]))
    assert errs[0] > errs[1] > errs[2] > errs[3]


def test_reconstruct_explicit_and_skip_invalid():
    x = np.random.default_rng(5).standard_normal(120)
    L, k = 30, 8
    s = O.OracleSSA(x, L).decompose(k, 6, 3, O.philox_gaussian(1, 0, 91, 14))
    U, V, sig = s.factors()
    for g in ([0], [2, 5], [1, 3, 7]):
        ref = diag_average(sum(sig[i] * np.outer(U[:, i], V[:, i]) for i in g))
        assert np.abs(s.reconstruct(g) - ref).max() < 1e-13 * np.abs(x).max() * 10
    # invalid indices are skipped silently (MKL/ssa_opt.h:2347-2348)
    assert np.array_equal(s.reconstruct([0, -1, 99]), s.reconstruct([0]))
    with pytest.raises(RuntimeError):
        s.reconstruct([])


def test_full_rank_reconstruction_is_identity():
    x = np.random.default_rng(6).standard_normal(64)
    L = 20
    s = O.OracleSSA(x, L).decompose(20, 4, 0, O.philox_gaussian(2, 0, 45, 20))
    assert np.abs(s.reconstruct(list(range(20))) - x).max() < 1e-12


def test_wcorr_definition():
    x = workloads.single(300, seed=3)
    L, k = 75, 10
    s = O.OracleSSA(x, L).decompose(k, 8, 2, O.philox_gaussian(3, 0, 226, 18))
    comps = np.stack([s.reconstruct([i]) for i in range(k)])
    Kd = len(x) - L + 1
    w = np.array([min(t + 1, L, Kd, len(x) - t) for t in range(len(x))], dtype=float)
    G = comps * np.sqrt(w)
    G /= np.linalg.norm(G, axis=1, keepdims=True)
    W = s.wcorr()
    assert np.abs(W - G @ G.T).max() < 1e-12
    assert np.allclose(np.diag(W), 1.0, atol=1e-13)
    assert np.array_equal(W, W.T)


def test_lrf_exact_on_sinusoid():
    N, L = 200, 40
    omega = 2 * np.pi / 13.0
    t = np.arange(N + 30)
    full = np.sin(omega * t + 0.4)
    s = O.OracleSSA(full[:N], L).decompose(2, 6, 2, O.philox_gaussian(4, 0, N - L + 1, 8))
    rc, R, nu = s.lrf([0, 1])
    assert rc == 0 and 0 <= nu < 1
    f = s.forecast([0, 1], 30)
    assert np.abs(f - full[N:]).max() < 1e-9
    ff = s.forecast_full([0, 1], 30)
    assert np.abs(ff[N:] - f).max() == 0.0
    # the recurrence itself: every window of the series satisfies sum R_j x_{t+j} = x_{t+L-1}
    win = np.lib.stride_tricks.sliding_window_view(full[:N], L)
    assert np.abs(win[:, :-1] @ R - win[:, -1]).max() < 1e-9


def test_lrf_verticality_failure():
    # a component whose eigenvector ends in 1 (e_L direction) is not forecastable
    x = np.zeros(50)
    x[-1] = 1.0
    s = O.OracleSSA(x, 10).decompose(1, 4, 1, O.philox_gaussian(5, 0, 41, 5))
    rc, _, nu = s.lrf([0])
    assert rc == -1 and nu >= 1 - 1e-10


def test_vforecast_exact_on_sinusoid_and_equal_to_lrf():
    """V-forecast (one projection per eigenvector and step) is the same
    recurrence as the LRF forecast: on a sinusoid it continues the series
    exactly, and on a noisy series it agrees with the R-forecast to rounding."""
    N, L = 200, 40
    omega = 2 * np.pi / 13.0
    t = np.arange(N + 30)
    full = np.sin(omega * t + 0.4)
    s = O.OracleSSA(full[:N], L).decompose(2, 6, 2, O.philox_gaussian(4, 0, N - L + 1, 8))
    v = s.vforecast([0, 1], 30)
    assert np.abs(v - full[N:]).max() < 1e-9
    x = workloads.single(600, seed=3)
    s2 = O.OracleSSA(x, 150).decompose(12, 8, 2, O.philox_gaussian(7, 0, 451, 20))
    g = [0, 1, 2, 3, 4]
    v2, r2 = s2.vforecast(g, 40), s2.forecast(g, 40)
    assert np.abs(v2 - r2).max() < 1e-10 * np.abs(r2).max()
    # the fast form continues the caller's base: given the reconstruction it is the plain form
    rec = s2.reconstruct(g)
    assert np.array_equal(s2.vforecast(g, 40, base=rec), v2)
    # and given another base it is the R-recurrence on that base
    base = x[:500]
    rc, R, _ = s2.lrf(g)
    buf = list(base[-149:])
    for _ in range(10):
        buf.append(float(np.dot(R, buf[-149:])))
    assert np.abs(s2.vforecast(g, 10, base=base) - np.array(buf[149:])).max() < 1e-10 * np.abs(base).max()


def test_vforecast_rejects_bad_input():
    x = workloads.single(300, seed=4)
    s = O.OracleSSA(x, 60).decompose(6, 6, 1, O.philox_gaussian(8, 0, 241, 12))
    with pytest.raises(RuntimeError):
        s.vforecast([0, 6], 5)             # index beyond the computed components
    with pytest.raises(RuntimeError):
        s.vforecast([0], 5, base=x[:58])   # base shorter than L - 1
    z = np.zeros(50)
    z[-1] = 1.0
    sv = O.OracleSSA(z, 10).decompose(1, 4, 1, O.philox_gaussian(5, 0, 41, 5))
    with pytest.raises(RuntimeError):
        sv.vforecast([0], 3)               # verticality >= 1


def test_mssa_matvec_and_single_channel():
    X = workloads.mssa_channels(M=3, N=120, seed=5)
    L = 30
    m = O.OracleMSSA(X, L)
    K = 120 - L + 1
    Hs = np.vstack([hankel(X[c], L) for c in range(3)])
    v = np.random.default_rng(1).standard_normal(K)
    u = np.random.default_rng(2).standard_normal(3 * L)
    assert np.abs(m.matvec(v) - Hs @ v).max() < 1e-12 * np.abs(X).max() * K
    assert np.abs(m.matvec_T(u) - Hs.T @ u).max() < 1e-12 * np.abs(X).max() * K
    om = O.philox_gaussian(6, 0, K, 14)
    m.decompose(6, 8, 3, om)
    sv = np.linalg.svd(Hs, compute_uv=False)
    assert np.allclose(m.factors()[2][:2], sv[:2], rtol=1e-10)
    # M = 1 is plain SSA (MKL/test/ssa_opt_test.c:1857-1892)
    m1 = O.OracleMSSA(X[:1], L).decompose(6, 8, 3, om)
    s1 = O.OracleSSA(X[0], L).decompose(6, 8, 3, om)
    assert np.allclose(m1.factors()[2], s1.sigma, rtol=1e-12)
    assert np.abs(m1.reconstruct(0, [0, 1]) - s1.reconstruct([0, 1])).max() < 1e-11
    c = m.contributions()
    assert np.allclose(c.sum(axis=0), 1.0, atol=1e-12)


def test_mssa_reconstruct_per_channel_explicit():
    """orc_mssa_reconstruct for M > 1 (MKL/ssa_opt.h:4212-4242): channel m of
    a group is the anti-diagonal average of sum_g sigma_g u_g^(m) v_g^T, where
    u_g^(m) = rows [m L, (m+1) L) of the stacked eigenvector; built here from
    the oracle's own factors with numpy, so a wrong channel offset or a
    transposed factor fails.  With every component of the stacked matrix
    (b = min(M L, K), an exact range) the channels sum back to the input."""
    M, N, L = 3, 60, 12
    X = workloads.mssa_channels(M=M, N=N, seed=7)
    K = N - L + 1
    full = min(M * L, K)
    m = O.OracleMSSA(X, L).decompose(full, 8, 1, O.philox_gaussian(8, 0, K, full))
    U, V, s = m.factors()
    Hs = np.vstack([hankel(X[c], L) for c in range(M)])
    assert np.allclose(s, np.linalg.svd(Hs, compute_uv=False)[:full], rtol=1e-10)
    for g in ([0], [1, 2], [0, 3, 5], list(range(full))):
        for ch in range(M):
            Xm = (U[ch * L:(ch + 1) * L, g] * s[g]) @ V[:, g].T
            assert np.abs(m.reconstruct(ch, g) - diag_average(Xm)).max() < 1e-11 * np.abs(X).max(), (g, ch)
    for ch in range(M):
        assert np.abs(m.reconstruct(ch, list(range(full))) - X[ch]).max() < 1e-10 * np.abs(X).max()
    # channels differ: the explicit average of another channel's rows does not match
    Xo = (U[L:2 * L, [0]] * s[[0]]) @ V[:, [0]].T
    assert np.abs(m.reconstruct(0, [0]) - diag_average(Xo)).max() > 1e-3 * np.abs(X).max()


# ---------------------------------------------------------------------------
# the reference's own dependency-free build (compiled from MKL/ssa_opt_ref.h)
# ---------------------------------------------------------------------------

needs_ref = pytest.mark.skipif(not O.ref_available(), reason="oracle/_ref not built")


@needs_ref
def test_against_reference_sequential_build():
    """ssa_opt_ref_decompose (power iteration, MKL/ssa_opt_ref.h) against the
    oracle at q = 8.  Separated components agree to rounding; the sine pair
    [1,2] is limited by the reference's own power-iteration convergence test
    (a per-step change below 1e-12), so it is compared at 1e-5."""
    x = workloads.single(1000, seed=1)
    L = 250
    r = O.RefSSA(x, L)
    r._l.ssa_opt_ref_decompose.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
    assert r._l.ssa_opt_ref_decompose(ctypes.byref(r.s), 6, 1000) == 0
    r.k = r.s.n_components
    s = O.OracleSSA(x, L).decompose(6, 10, 8, O.philox_gaussian(5, 0, 751, 16))
    rel = np.abs(s.sigma - r.sigma) / r.sigma
    assert rel[[0, 3, 4, 5]].max() < 1e-12
    assert rel[[1, 2]].max() < 1e-8
    rms = np.sqrt(np.mean(x ** 2))
    assert np.abs(s.reconstruct([0]) - r.reconstruct([0])).max() / rms < 1e-10
    assert np.abs(s.reconstruct([1, 2]) - r.reconstruct([1, 2])).max() / rms < 1e-5
    rc, fr = r.forecast([0, 1, 2], 20)
    assert rc == 0
    assert np.abs(fr - s.forecast_full([0, 1, 2], 20)).max() / rms < 1e-5


@needs_ref
def test_reference_randomized_qr_defect():
    """Documents why the reference's own randomized path cannot pin ours: its
    Householder QR (MKL/ssa_opt_ref.h:441-519) uses tau = beta/alpha, the
    wrong sign, so Q is not orthonormal and sigma is meaningless."""
    x = workloads.single(1000, seed=1)
    r = O.RefSSA(x, 250).decompose_randomized(10, 10)
    sv = np.linalg.svd(hankel(x, 250), compute_uv=False)
    assert abs(r.sigma[0] / sv[0] - 1.0) > 1e-3


def _sines(N, periods=(17.0, 41.0), amps=(1.0, 0.6), noise=0.0, seed=0):
    t = np.arange(N)
    x = sum(a * np.sin(2 * np.pi * t / p + 0.3 * i) for i, (a, p) in enumerate(zip(amps, periods)))
    if noise:
        x = x + noise * np.random.default_rng(seed).standard_normal(N)
    return x


def test_cadzow_single_step_is_one_reconstruction():
    """max_iter = 1 is one rank-r reconstruction with the Philox matrix of a
    fresh context (seed 42, counter 0), b = min(r + 8, L, K), q = 2."""
    x = _sines(300, noise=0.3, seed=1)
    L, r = 60, 4
    out, it, diff, conv = O.cadzow(x, L, r, max_iter=1, tol=0.0)
    K = 300 - L + 1
    ref = O.OracleSSA(x, L).decompose(r, 8, 2, O.philox_gaussian(42, 0, K, r + 8)).reconstruct(list(range(r)))
    assert it == 1 and not conv
    assert np.array_equal(out, ref)
    assert diff == pytest.approx(np.linalg.norm(ref - x) / np.linalg.norm(x), rel=1e-12)


def test_cadzow_fixed_point_and_rank_reduction():
    # a rank-4 signal (two sinusoids) is its own rank-4 Hankel projection
    x = _sines(400)
    out, it, diff, conv = O.cadzow(x, 80, 4, max_iter=5, tol=1e-9)
    assert conv and it <= 2 and np.abs(out - x).max() < 1e-9
    # on noisy data the iterations approach a rank-4 trajectory matrix
    y = _sines(400, noise=0.5, seed=2)
    out, it, diff, conv = O.cadzow(y, 80, 4, max_iter=60, tol=1e-14)
    s_in = np.linalg.svd(hankel(y, 80), compute_uv=False)
    s_out = np.linalg.svd(hankel(out, 80), compute_uv=False)
    assert s_out[4] / s_out[0] < 0.05 * s_in[4] / s_in[0]


def test_cadzow_inplace_converges_on_finite_rank():
    x = _sines(400)
    s = O.OracleSSA(x, 80)
    it, diff, conv = s.cadzow_inplace(4, 5, 1e-9)
    assert conv and it == 2 and diff < 1e-9
    assert np.abs(s.reconstruct([0, 1, 2, 3]) - x).max() < 1e-9


def test_gapfill_recovers_sinusoids():
    x = _sines(500)
    y = x.copy()
    y[200:215] = np.nan
    y[400] = np.nan
    filled, it, diff, conv, ng = O.gapfill(y, 80, 4, max_iter=200, tol=1e-12)
    assert ng == 16 and conv
    assert np.abs(filled - x).max() < 1e-6
    # the initial guess is the straight line between the neighbours
    z = x.copy()
    z[100:104] = np.nan
    lin, it1, _, _, _ = O.gapfill(z, 80, 4, max_iter=0, tol=0.0)
    assert it1 == 0 and np.allclose(lin[100:104], x[99] + (x[104] - x[99]) * np.arange(1, 5) / 5, atol=1e-15)
    # no gaps: untouched
    same, it0, _, conv0, ng0 = O.gapfill(x, 80, 4)
    assert ng0 == 0 and it0 == 0 and conv0 and np.array_equal(same, x)


def test_gapfill_simple_exact_on_sinusoids():
    x = _sines(600)
    y = x.copy()
    y[300:320] = np.nan
    filled, ng = O.gapfill_simple(y, 60, 4)
    assert ng == 20
    assert np.abs(filled - x).max() < 1e-8
    # a gap at the start has no left segment: the backcast alone fills it
    z = x.copy()
    z[0:10] = np.nan
    f0, _ = O.gapfill_simple(z, 60, 4)
    assert np.abs(f0 - x).max() < 1e-8
