"""The component spectra cache (ssa_opt_cache_ffts, MKL/ssa_opt.h:2110-2323):
reconstructions (MKL/ssa_opt.h:2338-2356) and the fast W-correlation
(MKL/ssa_opt.h:2492-2548) read the cached chunk spectra of U and V instead of
transforming the factors again.

Tolerances, same injected test matrix on both sides:
  cached vs uncached: plain 1e-12 of RMS(x) (W 1e-12); mixed 1e-4 of RMS(x), the
      mixed form's pointwise bound (the two paths round differently in FP32), W 1e-5
  against the oracle: plain 1e-9 of RMS(x), W 1e-10; mixed: every group 1e-4 of
      RMS(x) pointwise, and 1e-5 relative (2-norm) for the groups whose span is
      well separated ([0], [0, 1, 2], all k: a single component inside a cluster
      of near-equal singular values is not), W 1e-3 (noise-level components'
      W entries carry the FP32 rounding of their small reconstructions)
"""
import numpy as np
import pytest

import oracle as O
import workloads

pytestmark = pytest.mark.gpu


def _ssa():
    import ssa_amd
    return ssa_amd


def rms(x):
    return float(np.sqrt(np.mean(np.asarray(x) ** 2)))


def rel2(a, b):
    return float(np.linalg.norm(a - b) / np.linalg.norm(b))


def _decomposed(N, L, k, prec):
    S = _ssa()
    x = workloads.single(N, seed=N + 1)
    K = N - L + 1
    om = O.philox_gaussian(77 + N, 0, K, min(k + 10, L, K))
    s = S.SSA(x, L, q=2, precision=prec)
    s.prepare(k, 10)
    s.inject_omega(om)
    s.decompose(k, oversampling=10)
    return x, s, om


# (N, L, k, precision): one chunk each way, several column chunks, several row chunks
CASES = [(1000, 250, 50, 0), (20000, 5000, 50, 0), (9000, 7000, 20, 0), (20000, 5000, 50, 1), (4000, 1000, 30, 1)]


@pytest.mark.parametrize("N,L,k,prec", CASES)
def test_cached_matches_uncached_and_oracle(N, L, k, prec):
    x, s, om = _decomposed(N, L, k, prec)
    groups = [[i] for i in range(k)] + [[0, 1, 2], list(range(k)), [1, 99, -3, 4]]
    slow = s.reconstruct_many(groups)
    W_slow = s.wcorr_matrix()
    assert not s.fft_cached
    s.cache_ffts()
    assert s.fft_cached
    fast = s.reconstruct_many(groups)
    W_fast = s.wcorr_matrix_fast()
    # the slow W-correlation keeps forming the transforms itself
    assert np.array_equal(s.wcorr_matrix(), W_slow)
    r = rms(x)
    d_rec = float(np.abs(fast - slow).max() / r)
    d_w = float(np.abs(W_fast - W_slow).max())
    orc = O.OracleSSA(x, L).decompose(k, 10, 2, om)
    ref = np.array([orc.reconstruct(g) for g in groups])
    W_orc = orc.wcorr()
    print(f"cache N={N} L={L} k={k} prec={prec} plan={s.plan_info()}: cached vs uncached {d_rec:.1e} "
          f"(W {d_w:.1e}); vs oracle: recon {float(np.abs(fast - ref).max() / r):.1e}, "
          f"W {float(np.abs(W_fast - W_orc).max()):.1e}")
    if prec == 0:
        assert d_rec < 1e-12 and d_w < 1e-12
        assert np.abs(fast - ref).max() / r < 1e-9
        assert np.abs(W_fast - W_orc).max() < 1e-10
    else:
        assert d_rec < 1e-4 and d_w < 1e-5
        assert np.abs(fast - ref).max() / r < 1e-4
        assert np.abs(W_fast - W_orc).max() < 1e-3
        for g, a, b in zip(groups, fast, ref):
            if g in ([0], [0, 1, 2], list(range(k))):
                assert rel2(a, b) < 1e-5, g
    # a single reconstruction reads the cache too (its piece shares an inverse
    # transform with a different partner than in the launch above)
    assert np.abs(s.reconstruct([0, 1, 2]) - fast[k]).max() / r < (1e-12 if prec == 0 else 1e-4)


def test_cache_lifetime():
    """The cache needs a decomposition, and update_signal, every decomposition
    and free_cached_ffts drop it; afterwards the uncached path gives the same
    bits as before the cache existed."""
    S = _ssa()
    x, s, om = _decomposed(2000, 500, 12, 0)
    before = s.reconstruct_many([[0], [1, 2]])
    s.cache_ffts()
    s.free_cached_ffts()
    assert not s.fft_cached
    assert np.array_equal(s.reconstruct_many([[0], [1, 2]]), before)
    s.cache_ffts()
    s.update_signal(x)
    assert not s.fft_cached
    s.inject_omega(om)
    s.decompose(12, oversampling=10)
    assert np.array_equal(s.reconstruct_many([[0], [1, 2]]), before)
    s.cache_ffts()
    s.inject_omega(om)
    s.decompose(12, oversampling=10)          # a decomposition drops the cache
    assert not s.fft_cached
    assert np.array_equal(s.reconstruct_many([[0], [1, 2]]), before)
    fresh = S.SSA(x, 500)
    assert fresh.lib.ssa_opt_cache_ffts(fresh.ref) == -1   # MKL/ssa_opt.h:2112-2113: needs a decomposition
    # without a cache the fast W-correlation is the slow one
    assert np.array_equal(s.wcorr_matrix_fast(), s.wcorr_matrix())
