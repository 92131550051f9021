"""The rest of the reference's analysis ABI on the device path, against the
FP64 oracle: the subspace-iteration entry points (ssa_opt_decompose,
_decompose_block, _extend; MKL/ssa_opt.h:1560-2056), W-correlation of one
pair, periodic-pair detection, component statistics, trend and noise
(MKL/ssa_opt.h:2550-2756, :3086-3110).

The subspace entry points run the device subspace iteration with
kp = min(max(k + 8, block), L, K) columns and q = clamp(max_iter / 4, 2, 30)
iterations on the context's Philox test matrix (counter = decompositions so
far), so the oracle restates them with the same matrix: singular values
within 1e-10 relative, reconstructions within 1e-9 of RMS(x)."""
import ctypes

import numpy as np
import pytest

import oracle as O
import workloads

pytestmark = pytest.mark.gpu

SEED = 77


def _ssa():
    import ssa_amd
    return ssa_amd


def rms(x):
    return float(np.sqrt(np.mean(np.asarray(x) ** 2)))


def _sv(s, n=80):
    out = np.zeros(n)
    m = s.lib.ssa_opt_get_singular_values(s.ref, out.ctypes.data_as(ctypes.POINTER(ctypes.c_double)), n)
    return out[:m]


def _oracle(x, L, k, kp, q, call):
    K = len(x) - L + 1
    return O.OracleSSA(x, L).decompose(k, kp - k, q, O.philox_gaussian(SEED, call, K, kp))


def test_subspace_entry_points_match_oracle():
    S = _ssa()
    x = workloads.single(1200, seed=21)
    L = 300
    r = rms(x)
    s = S.SSA(x, L, seed=SEED)
    # sequential power iteration entry point: k = 8, kp = 16, q = 40 / 4 = 10
    s.decompose(8, method="sequential", max_iter=40)
    o = _oracle(x, L, 8, 16, 10, call=0)
    assert np.max(np.abs(_sv(s) - o.sigma) / o.sigma) < 1e-10
    for g in ([0], [1, 2], list(range(8))):
        assert np.abs(s.reconstruct(g) - o.reconstruct(g)).max() / r < 1e-9, g
    # block entry point: kp = max(k + 8, block) = 20, q = max(8 / 4, 2) = 2
    s.decompose(6, method="block", block_size=20, max_iter=8)
    o = _oracle(x, L, 6, 20, 2, call=1)
    assert np.max(np.abs(_sv(s) - o.sigma) / o.sigma) < 1e-10
    # extend by 4: a sequential decomposition of k = 10 (kp = 18, q = 10)
    assert s.lib.ssa_opt_extend(s.ref, 4, 40) == 0
    o = _oracle(x, L, 10, 18, 10, call=2)
    sv = _sv(s)
    assert len(sv) == 10
    assert np.max(np.abs(sv - o.sigma) / o.sigma) < 1e-10
    assert np.abs(s.reconstruct([3, 4]) - o.reconstruct([3, 4])).max() / r < 1e-9
    # extend requires a decomposition; invalid arguments are refused
    assert s.lib.ssa_opt_extend(s.ref, 0, 40) == -1


def _pair_signal(N=2000):
    t = np.arange(N)
    rng = np.random.default_rng(5)
    return 0.002 * t + 3.0 * np.sin(2 * np.pi * t / 50) + 1.5 * np.sin(2 * np.pi * t / 12 + 0.4) \
        + 0.05 * rng.standard_normal(N)


def test_wcorr_pair_trend_noise_and_stats():
    S = _ssa()
    x = _pair_signal()
    L, k, p, q = 400, 12, 10, 2
    K = len(x) - L + 1
    om = O.philox_gaussian(1234, 0, K, k + p)
    s = S.SSA(x, L, q=q)
    s.prepare(k, p)
    s.inject_omega(om)
    s.decompose(k, oversampling=p)
    o = O.OracleSSA(x, L).decompose(k, p, q, om)
    W = o.wcorr()
    for i, j in ((0, 1), (1, 2), (3, 4), (2, 7), (5, 5)):
        assert abs(s.wcorr(i, j) - W[i, j]) < 1e-10, (i, j)
    assert s.wcorr(0, k) == 0.0          # out of range: 0.0, as the reference
    r = rms(x)
    assert np.abs(s.get_trend() - o.reconstruct([0])).max() / r < 1e-9
    assert np.abs(s.get_noise(5) - o.reconstruct(list(range(5, k)))).max() / r < 1e-9

    class Stats(ctypes.Structure):
        _fields_ = [("n", ctypes.c_int), ("singular_values", ctypes.POINTER(ctypes.c_double)),
                    ("log_sv", ctypes.POINTER(ctypes.c_double)), ("gaps", ctypes.POINTER(ctypes.c_double)),
                    ("cumulative_var", ctypes.POINTER(ctypes.c_double)),
                    ("second_diff", ctypes.POINTER(ctypes.c_double)), ("suggested_signal", ctypes.c_int),
                    ("gap_threshold", ctypes.c_double)]

    st = Stats()
    assert s.lib.ssa_opt_component_stats(s.ref, ctypes.byref(st)) == 0
    try:
        so = o.sigma
        n = st.n
        assert n == k
        gaps = np.array(st.gaps[:n - 1])
        assert np.allclose(np.array(st.singular_values[:n]), so, rtol=1e-10, atol=0)
        assert np.allclose(gaps, so[:-1] / so[1:], rtol=1e-9, atol=0)
        cv = np.array(st.cumulative_var[:n])
        assert np.allclose(cv, np.cumsum(so ** 2) / np.sum(so ** 2), rtol=1e-9, atol=0) and cv[-1] == pytest.approx(1.0)
        # the largest singular-value ratio separates the linear trend (two
        # components) and the two sinusoids (two each) from the noise floor
        assert st.suggested_signal == int(np.argmax(so[:-1] / so[1:])) + 1 == 6
        assert st.gap_threshold == pytest.approx(float(np.max(so[:-1] / so[1:])), rel=1e-9)
        ls = np.log(so)
        sd = np.array(st.second_diff[:n])
        assert sd[0] == sd[-1] == 0.0
        assert np.allclose(sd[1:-1], ls[:-2] - 2 * ls[1:-1] + ls[2:], rtol=0, atol=1e-9)
    finally:
        s.lib.ssa_opt_component_stats_free(ctypes.byref(st))


def test_periodic_pairs_on_two_sinusoids():
    """Known answer: a linear trend plus sinusoids of periods 50 and 12
    decompose into the trend (components 0 and 5), the pair (1, 2) of period
    50 and the pair (3, 4) of period 12; the device's
    pairs also follow the reference's rule (MKL/ssa_opt.h:2690-2756) applied
    to the oracle's singular values and W-correlations."""
    S = _ssa()
    x = _pair_signal()
    L, k, p, q = 400, 12, 10, 2
    K = len(x) - L + 1
    om = O.philox_gaussian(4321, 0, K, k + p)
    s = S.SSA(x, L, q=q)
    s.prepare(k, p)
    s.inject_omega(om)
    s.decompose(k, oversampling=p)
    pairs = s.find_periodic_pairs(max_pairs=4, sv_tol=0.1, wcorr_thresh=0.5)
    assert pairs[:2] == [(1, 2), (3, 4)]
    o = O.OracleSSA(x, L).decompose(k, p, q, om)
    so, W = o.sigma, o.wcorr()
    want, used = [], set()
    for i in range(k - 1):
        if i in used or len(want) == 4:
            continue
        for j in range(i + 1, k):
            if j in used:
                continue
            if abs(1 - so[j] / so[i]) > 0.1 or abs(W[i, j]) < 0.5:
                continue
            want.append((i, j))
            used.update((i, j))
            break
    assert pairs == want
    assert s.find_periodic_pairs(max_pairs=1) == [(1, 2)]
