"""ESPRIT core (host FP64: Householder least squares, Hessenberg reduction and
Francis double-shift QR) of ssa_opt_parestimate (MKL/ssa_opt.h:4304-4558),
through the C ABI without a device: poles against numpy's lstsq + eigvals on
random subspaces, and exactly the signal poles on finite-rank signals."""
import numpy as np
import pytest


def _S():
    import ssa_amd
    return ssa_amd


def numpy_poles(U):
    Z = np.linalg.lstsq(U[:-1], U[1:], rcond=None)[0]
    return np.linalg.eigvals(Z)


def matched(a, b):
    """max distance after greedy nearest matching of two pole sets"""
    b = list(b)
    worst = 0.0
    for z in a:
        j = int(np.argmin([abs(z - w) for w in b]))
        worst = max(worst, abs(z - b[j]) / max(1.0, abs(b[j])))
        b.pop(j)
    return worst


@pytest.mark.parametrize("L,k", [(3, 1), (3, 2), (10, 3), (40, 5), (200, 20), (300, 60), (500, 80)])
def test_poles_match_numpy(L, k):
    rng = np.random.default_rng(L * 100 + k)
    U = np.linalg.qr(rng.standard_normal((L, k)))[0]
    p = _S().esprit_poles(U)
    assert len(p) == k
    assert matched(p, numpy_poles(U)) < 1e-10
    # conjugate pairs come out as exact conjugates
    assert matched(p, np.conj(p)) < 1e-12


def test_poles_of_finite_rank_signals():
    # sum of a damped sinusoid, an undamped one and a decaying exponential:
    # rank 5, poles rho e^{+-i w}, e^{+-i w2}, r
    N, L = 400, 120
    t = np.arange(N)
    x = (0.98 ** t) * np.cos(2 * np.pi * t / 17) + 0.7 * np.sin(2 * np.pi * t / 41 + 0.2) + 1.5 * 0.995 ** t
    H = np.lib.stride_tricks.sliding_window_view(x, N - L + 1)[:L]
    U = np.linalg.svd(H, full_matrices=False)[0][:, :5]
    p = _S().esprit_poles(U)
    want = [0.98 * np.exp(2j * np.pi / 17), 0.98 * np.exp(-2j * np.pi / 17), np.exp(2j * np.pi / 41),
            np.exp(-2j * np.pi / 41), 0.995]
    assert matched(p, want) < 1e-9


def test_rejects_degenerate_input():
    S = _S()
    with pytest.raises(RuntimeError):
        S.esprit_poles(np.ones((2, 1)))             # L < 3
    with pytest.raises(RuntimeError):
        S.esprit_poles(np.zeros((20, 2)))           # U_up rank deficient
