"""Cadzow iterations and gap filling on the device (MKL/ssa_opt.h:3147-3878)
against the oracle's restatement, plain FP64: every decomposition draws the
Philox matrix a fresh context would (seed 42, counter 0; the in-place variant
advances the counter), so both sides see the same test matrices.  Outputs
within 1e-9 of RMS(x), iteration counts equal."""
import numpy as np
import pytest

import oracle as O

pytestmark = pytest.mark.gpu


def _S():
    import ssa_amd
    return ssa_amd


def _sines(N, noise=0.0, seed=0):
    t = np.arange(N)
    x = np.sin(2 * np.pi * t / 17.0) + 0.6 * np.sin(2 * np.pi * t / 41.0 + 0.3)
    if noise:
        x = x + noise * np.random.default_rng(seed).standard_normal(N)
    return x


def rms(x):
    return float(np.sqrt(np.nanmean(np.asarray(x) ** 2)))


@pytest.mark.parametrize("N,L,rank", [(400, 80, 4), (400, 20, 4), (900, 300, 6)])
def test_cadzow_against_oracle(N, L, rank):
    S = _S()
    x = _sines(N, noise=0.3, seed=N + L)
    for max_iter in (1, 5):
        out, it, diff, conv = S.cadzow(x, L, rank, max_iter=max_iter, tol=0.0)
        o, oit, odiff, oconv = O.cadzow(x, L, rank, max_iter=max_iter, tol=0.0)
        assert it == oit == max_iter and conv == oconv
        assert np.abs(out - o).max() / rms(x) < 1e-9
        assert abs(diff - odiff) < 1e-9
    # converged runs stop at the same iteration
    out, it, diff, conv = S.cadzow(x, L, rank, max_iter=40, tol=1e-4)
    o, oit, odiff, oconv = O.cadzow(x, L, rank, max_iter=40, tol=1e-4)
    assert (it, conv) == (oit, oconv)
    assert np.abs(out - o).max() / rms(x) < 1e-9


def test_cadzow_weighted_and_errors():
    S = _S()
    x = _sines(400, noise=0.3, seed=3)
    c, *_ = S.cadzow(x, 80, 4, max_iter=3, tol=0.0)
    w, it, _, _ = S.cadzow(x, 80, 4, max_iter=3, tol=0.0, alpha=0.7)
    assert it == 3
    assert np.abs(w - (0.7 * c + 0.3 * x)).max() < 1e-12
    for bad in (dict(alpha=0.0), dict(alpha=1.5)):
        with pytest.raises(RuntimeError):
            S.cadzow(x, 80, 4, **bad)
    with pytest.raises(RuntimeError):
        S.cadzow(x, 400, 4)            # L > N - 1
    with pytest.raises(RuntimeError):
        S.cadzow(x, 80, 0)             # rank < 1


def test_cadzow_inplace_against_oracle():
    S = _S()
    x = _sines(500, noise=0.2, seed=4)
    dev = S.SSA(x, 100, q=2, seed=42)
    it, diff, conv = dev.cadzow_inplace(4, 4, 0.0)
    orc = O.OracleSSA(x, 100)
    oit, odiff, oconv = orc.cadzow_inplace(4, 4, 0.0, counter0=0)
    assert (it, conv) == (oit, oconv) == (4, False)
    assert abs(diff - odiff) < 1e-9
    # the context keeps the last decomposition
    assert np.abs(dev.reconstruct([0, 1, 2, 3]) - orc.reconstruct([0, 1, 2, 3])).max() / rms(x) < 1e-9


def test_gapfill_against_oracle():
    S = _S()
    x = _sines(600, noise=0.1, seed=5)
    y = x.copy()
    y[250:270] = np.nan
    y[[10, 11, 500]] = np.nan
    for max_iter, tol in ((3, 0.0), (100, 1e-5)):
        d = S.gapfill(y, 90, 4, max_iter=max_iter, tol=tol)
        o = O.gapfill(y, 90, 4, max_iter=max_iter, tol=tol)
        assert d[1:3] == pytest.approx(o[1:3], abs=1e-9) and d[3:] == o[3:]
        assert d[4] == 23
        assert np.abs(d[0] - o[0]).max() / rms(x) < 1e-9
        assert np.array_equal(np.isnan(d[0]), np.zeros(600, bool))
    with pytest.raises(RuntimeError):
        S.gapfill(np.where(np.arange(600) < 520, np.nan, x), 90, 4)   # too few observed values


def test_gapfill_simple_against_oracle():
    """One gap per call: the reference backcasts from the whole tail of the
    series, which still holds any later gap's NaNs (MKL/ssa_opt.h:3770-3778)."""
    S = _S()
    x = _sines(700, noise=0.05, seed=6)
    for lo, hi in ((300, 330), (0, 5), (690, 700)):
        y = x.copy()
        y[lo:hi] = np.nan
        d, it, _, conv, ng = S.gapfill(y, 70, 4, simple=True)
        o, ong = O.gapfill_simple(y, 70, 4)
        assert (it, conv, ng) == (1, True, hi - lo) and ong == hi - lo
        assert np.abs(d - o).max() / rms(x) < 1e-9
        assert np.abs(d[lo:hi] - x[lo:hi]).max() < 0.5   # the forecasts follow the sinusoids through the noise
