"""The structural checks of the reference's C test suite
(MKL/test/ssa_opt_test.c), run on the device path through the same entry
points: the shapes, signals, decomposition methods and thresholds of each
reference test, asserted on this library's results.  The oracle parity of
these paths is pinned elsewhere (test_gpu_parity.py, test_gpu_analysis.py);
these tests show that a caller of the reference's C API sees the behaviour
its own tests expect."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _ssa():
    import ssa_amd
    return ssa_amd


def _corr(a, b):
    a = np.asarray(a) - np.mean(a)
    b = np.asarray(b) - np.mean(b)
    return float(a @ b / np.sqrt((a @ a) * (b @ b)))


@pytest.mark.parametrize("method", ["sequential", "randomized"])
def test_constant_signal(method):
    """ssa_opt_test.c:1289-1322: a constant series is rank 1: the first
    eigenvalue dominates (ratio > 1000) and component 0 reconstructs the
    constant to 0.01."""
    S = _ssa()
    x = np.full(200, 42.0)
    s = S.SSA(x, 50).decompose(5, method=method)
    lam = s.eigenvalues
    assert lam[0] / (lam[1] + 1e-15) > 1000
    assert np.max(np.abs(s.reconstruct([0]) - 42.0)) < 0.01


@pytest.mark.parametrize("method", ["sequential", "randomized"])
def test_linear_trend_only(method):
    """ssa_opt_test.c:1773-1807: a linear trend is rank 2: the top two
    eigenvalues hold over 100x the next three, and their reconstruction
    correlates with the series above 0.999."""
    S = _ssa()
    t = np.arange(200)
    x = 3.0 + 0.5 * t
    s = S.SSA(x, 50).decompose(5, method=method)
    lam = s.eigenvalues
    assert (lam[0] + lam[1]) / (lam[2:5].sum() + 1e-10) > 100
    assert _corr(s.reconstruct([0, 1]), x) > 0.999


def _test_signal(N, seed):
    """The reference suite's test signal (ssa_opt_test.c:128-139): trend
    100 + 50 t/N, cycles of period 100 (amplitude 20) and 25 (amplitude 10),
    Gaussian noise of deviation 5 (here from numpy)."""
    i = np.arange(N)
    rng = np.random.default_rng(seed)
    return (100.0 + 50.0 * i / N + 20.0 * np.sin(2 * np.pi * i / 100.0) + 10.0 * np.sin(2 * np.pi * i / 25.0)
            + 5.0 * rng.standard_normal(N))


@pytest.mark.parametrize("noise", [True, False])
def test_randomized_decomposition_flags_and_order(noise):
    """ssa_opt_test.c:250-280: a randomized decomposition (N = 500, L = 100,
    k = 10, p = 8) returns 0, sets the decomposed and prepared flags and
    n_components, and gives singular values sigma + 1e-10 > 0 in decreasing
    order (sigma_i + 1e-10 > sigma_i+1).  Without the noise the series has
    rank 6: the four directions beyond it hold rounding only, and their sigma
    is at the rounding level or exactly 0 (a Cholesky pivot below the
    dead-column floor), which the reference's assertions allow."""
    S = _ssa()
    x = _test_signal(500, 12345)
    if not noise:
        x = x - 5.0 * np.random.default_rng(12345).standard_normal(500)
    s = S.SSA(x, 100)
    s.prepare(10, 8)
    s.decompose(10, oversampling=8)
    assert s.ctx.decomposed and s.ctx.prepared and s.n_components == 10
    sv = s.singular_values
    assert np.all(sv + 1e-10 > 0)
    assert np.all(sv[:-1] + 1e-10 > sv[1:])
    if noise:
        assert np.all(sv > 0)
    else:
        assert np.all(sv[:6] > 1e-6 * sv[0]) and np.all(sv[6:] <= 1e-14 * sv[0])


def test_randomized_determinism():
    """ssa_opt_test.c:1735-1771: two fresh contexts on the same series give
    the same singular values (the reference allows 1 % of sigma_0; here the
    test matrix comes from the same Philox key, so they are bit-identical)."""
    S = _ssa()
    t = np.arange(300)
    x = np.sin(0.1 * t) + 0.5 * np.cos(0.23 * t)
    a = S.SSA(x, 75)
    a.prepare(10, 8)
    a.decompose(10, oversampling=8)
    b = S.SSA(x, 75)
    b.prepare(10, 8)
    b.decompose(10, oversampling=8)
    assert np.array_equal(a.singular_values, b.singular_values)


def test_invalid_and_extreme_windows():
    """ssa_opt_test.c:1629-1666: L = 1 is refused; L = N - 1 (K = 2) is
    accepted and decomposes."""
    S = _ssa()
    x = np.sin(0.1 * np.arange(100))
    with pytest.raises(ValueError):
        S.SSA(x, 1)
    s = S.SSA(x, 99)
    assert s.K == 2
    s.decompose(1, oversampling=1)
    assert s.n_components == 1 and s.singular_values[0] > 0


def test_k_exceeds_rank():
    """ssa_opt_test.c:1668-1688: asking for more components than
    min(L, K) = 20 is clamped (or refused); a decomposition never reports
    more components than the rank allows."""
    S = _ssa()
    x = np.sin(0.1 * np.arange(100))
    s = S.SSA(x, 20)
    lib, ctx = s.lib, s.ref
    rc = lib.ssa_opt_decompose(ctx, 25, 100)
    assert rc != 0 or s.ctx.n_components <= 20
    assert lib.ssa_opt_prepare(ctx, 25, 8) == 0
    rc = lib.ssa_opt_decompose_randomized(ctx, 25, 8)
    assert rc != 0 or s.ctx.n_components <= 20


def test_small_signal_and_large_window():
    """ssa_opt_test.c:1368-1412: the minimum viable series (N = 20, L = 5,
    k = 3) decomposes to exactly k components; a window near N/2 (N = 200,
    L = 99) gives K = N - L + 1 and decomposes."""
    S = _ssa()
    s = S.SSA(np.sin(0.5 * np.arange(20)), 5).decompose(3, method="sequential")
    assert s.n_components == 3
    t = np.arange(200)
    s = S.SSA(np.sin(0.1 * t) + 0.5 * np.cos(0.3 * t), 99)
    assert s.K == 102
    s.decompose(10, method="sequential")
    assert s.n_components == 10


def test_empty_group_reconstruction():
    """ssa_opt_test.c:1894-1917: a reconstruction of an empty group is
    handled (here refused with -1, as n_group < 1 is in the reference's
    ssa_opt_reconstruct), and the context stays usable."""
    S = _ssa()
    x = np.sin(0.1 * np.arange(200))
    s = S.SSA(x, 50).decompose(5, method="sequential")
    out = np.zeros(200)
    rc = s.lib.ssa_opt_reconstruct(s.ref, None, 0, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)))
    assert rc == -1
    assert np.all(np.isfinite(s.reconstruct([0, 1])))


def test_wcorr_matrix_properties():
    """ssa_opt_test.c:411-454: the W-correlation matrix has a unit diagonal
    (within 0.01), is symmetric (within 1e-10), and a sine/cosine pair is
    strongly correlated (above 0.5)."""
    S = _ssa()
    t = np.arange(500)
    x = np.sin(2 * np.pi * t / 50) + 0.3 * np.sin(2 * np.pi * t / 11) + 0.05 * np.random.default_rng(3).standard_normal(500)
    s = S.SSA(x, 125)
    s.prepare(10, 8)
    s.decompose(10, oversampling=8)
    W = s.wcorr_matrix()
    assert np.max(np.abs(np.diag(W) - 1.0)) < 0.01
    assert np.max(np.abs(W - W.T)) < 1e-10
    assert abs(W[0, 1]) > 0.5


def test_forecast_rmse_and_long_horizon():
    """ssa_opt_test.c:2117-2175 and :1919-1955: the R-forecast of a sinusoid
    with a trend (period 40, N = 400, L = 100, components 0..2) has a
    10-step RMSE below 1.0 against the true continuation; a 500-step
    forecast of a pure sinusoid (N = 300, L = 75, components 0, 1) returns
    and stays bounded (below 10)."""
    S = _ssa()
    N, period = 400, 40.0
    t = np.arange(N)
    x = 10.0 * np.sin(2 * np.pi * t / period) + 0.01 * t
    s = S.SSA(x, 100).decompose(6, method="sequential")
    f = s.forecast([0, 1, 2], 10)
    tf = np.arange(N, N + 10)
    truth = 10.0 * np.sin(2 * np.pi * tf / period) + 0.01 * tf
    assert np.sqrt(np.mean((f - truth) ** 2)) < 1.0
    x2 = np.sin(2 * np.pi * np.arange(300) / 50.0)
    s2 = S.SSA(x2, 75).decompose(4, method="sequential")
    f2 = s2.forecast([0, 1], 500)
    assert f2.shape == (500,) and np.max(np.abs(f2)) < 10.0


def test_variance_partition():
    """ssa_opt_test.c:2342-2381: the explained-variance fractions of the
    components add up to that of the whole range and lie in [0, 1]."""
    S = _ssa()
    t = np.arange(600)
    x = np.sin(2 * np.pi * t / 30) + 0.2 * np.cos(2 * np.pi * t / 7) + 0.001 * t * t / 600
    s = S.SSA(x, 150)
    s.prepare(12, 8)
    s.decompose(12, oversampling=8)
    parts = [s.variance_explained(i, i) for i in range(12)]
    whole = s.variance_explained(0, 11)
    assert all(0.0 <= p <= 1.0 for p in parts)
    assert abs(sum(parts) - whole) < 1e-12
    assert 0.0 < whole <= 1.0 + 1e-12
