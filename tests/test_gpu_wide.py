"""Parity of the block-matvec instantiations the batched benchmark times.

A block-matvec launch of at least SSA_AMD_MV_WIDE_MIN items (default 2048)
takes the wide transform geometry: k_matvec<float, 13, true> (32 values per
thread, a radix-32 pass then two radix-16 passes) in the mixed form and
k_matvec<double, 10..12, true> (radix-16 passes at two waves per SIMD) in the
plain form.  Single series never reach that threshold, so these tests lower
it to 1 and compare the wide kernels with the FP64 oracle (MKL/ssa_opt.h:910-1004
for the block matvec), then run the benchmark's own 4096-series batch (both
forms, HIP-graph path, two branches) and compare series of every FFT-length
bucket with the oracle.

Tolerances (same injected or Philox test matrix on both sides):
  plain FP64: singular values 1e-10 relative, reconstructions 1e-9 of RMS(x)
  mixed:      singular values 1e-5 relative, reconstructions 1e-5 relative
              (2-norm) and 1e-4 of RMS(x) pointwise
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


def hankel(x, L):
    K = len(x) - L + 1
    return np.lib.stride_tricks.sliding_window_view(x, K)[:L]


# (precision, N, FFT length the planner picks for N, L = N / 4): every length
# whose wide instantiation differs from the narrow one
WIDE_CASES = [(1, 20000, 8192), (1, 12000, 8192), (0, 1500, 1024), (0, 3000, 2048), (0, 6000, 4096)]


@pytest.mark.parametrize("prec,N,n", WIDE_CASES)
def test_wide_matvec_bruteforce(monkeypatch, prec, N, n):
    """Single-vector H v and H^T u through the wide kernels against the
    explicit Hankel products."""
    S = _ssa()
    monkeypatch.setenv("SSA_AMD_MV_WIDE_MIN", "1")
    L = N // 4
    x = np.random.default_rng(N).standard_normal(N)
    s = S.SSA(x, L, precision=prec)
    assert s.plan_info()["n"] == n
    K = N - L + 1
    v = np.random.default_rng(1).standard_normal(K)
    u = np.random.default_rng(2).standard_normal(L)
    H = hankel(x, L)
    ry, rz = H @ v, H.T @ u
    tol = 1e-11 if prec == 0 else 1e-5
    assert np.abs(s.hankel_matvec(v) - ry).max() / np.abs(ry).max() < tol
    assert np.abs(s.hankel_matvec_T(u) - rz).max() / np.abs(rz).max() < tol


@pytest.mark.parametrize("prec,N,n", WIDE_CASES)
def test_wide_decomposition_parity(monkeypatch, prec, N, n):
    """Decomposition (q = 2, p = 10) and reconstructions with every block
    matvec on the wide kernels, against the oracle and against the same
    context on the narrow kernels."""
    S = _ssa()
    L, k, p = N // 4, 40, 10
    K = N - L + 1
    x = workloads.single(N, seed=N)
    om = O.philox_gaussian(500 + N, 0, K, k + p)

    def run():
        d = S.SSA(x, L, q=2, precision=prec)
        d.prepare(k, p)
        d.inject_omega(om)
        d.decompose(k, oversampling=p)
        return d

    narrow = run()
    monkeypatch.setenv("SSA_AMD_MV_WIDE_MIN", "1")
    wide = run()
    assert wide.plan_info()["n"] == n
    orc = O.OracleSSA(x, L).decompose(k, p, 2, om)
    so, r = orc.sigma, rms(x)
    groups = ([0], [1, 2], list(range(k)))
    if prec == 0:
        assert np.max(np.abs(wide.singular_values - so) / so) < 1e-10
        for g in groups:
            assert np.abs(wide.reconstruct(g) - orc.reconstruct(g)).max() / r < 1e-9, g
        # wide and narrow differ only in the transforms' rounding
        assert np.max(np.abs(wide.singular_values - narrow.singular_values) / so) < 1e-12
    else:
        assert np.max(np.abs(wide.singular_values - so) / so) < 1e-5
        for g in groups:
            a, b = wide.reconstruct(g), orc.reconstruct(g)
            assert rel2(a, b) < 1e-5, g
            assert np.abs(a - b).max() / r < 1e-4, g


def _batch_key(seed, i):
    return (seed ^ (0x9E3779B97F4A7C15 * (i + 1))) & ((1 << 64) - 1)


@pytest.mark.parametrize("prec", [1, 0])
def test_benchmark_batch_parity(prec):
    """The benchmark step itself: workloads.ragged_batch(4096, seed=4), one
    HIP-graph launch over two branches (the launches of the dominant bucket
    hold thousands of items, so they take the wide kernels).  Every series
    must report status 0 and finite output; three series of every FFT-length
    bucket, and the longest series, are compared with the oracle (their own
    Philox keys)."""
    S = _ssa()
    xs, Ns, Ls, ks = workloads.ragged_batch(count=4096, seed=4, dtype=np.float32)
    b = S.Batch(xs, Ls, ks, p=10, q=2, precision=prec, seed=4)
    b.run(use_graph=True)
    b.sync()
    outs, status = b.outputs()
    assert not np.any(status), np.flatnonzero(status)[:10]
    assert all(np.all(np.isfinite(o)) for o in outs)
    plans = [b.plan(i) for i in range(len(xs))]
    buckets = {}
    for i, pl in enumerate(plans):
        buckets.setdefault(pl["lgn"], []).append(i)
    pick = sorted({i for v in buckets.values() for i in v[:3]} | {int(np.argmax(Ns))})
    assert len(pick) >= 16 and 13 in buckets
    # the n = 8192 bucket's launches take the wide geometry at the default
    # threshold: every branch's forward launch holds one item per column pair
    pairs = sum((plans[i]["b"] + 1) // 2 for i in buckets[13])
    assert pairs // 2 >= 2048, pairs
    worst_s = worst_r = worst_p = 0.0
    for i in pick:
        x = np.asarray(xs[i], dtype=np.float64)
        L, k = int(Ls[i]), int(ks[i])
        K = len(x) - L + 1
        bw = min(k + 10, L, K)
        orc = O.OracleSSA(x, L).decompose(k, 10, 2, O.philox_gaussian(_batch_key(4, i), 0, K, bw))
        so = orc.sigma[:k]
        rec = orc.reconstruct(list(range(k)))
        worst_s = max(worst_s, float(np.max(np.abs(b.sigma(i)[:k] - so) / so)))
        worst_r = max(worst_r, rel2(outs[i], rec))
        worst_p = max(worst_p, float(np.abs(outs[i] - rec).max() / rms(x)))
    print(f"benchmark batch prec={prec}: {len(pick)} series over buckets {sorted(buckets)}: "
          f"sigma {worst_s:.2e}, recon rel2 {worst_r:.2e}, pointwise/RMS {worst_p:.2e}")
    if prec == 0:
        assert worst_s < 1e-10 and worst_p < 1e-9
    else:
        assert worst_s < 1e-5 and worst_r < 1e-5 and worst_p < 1e-4
