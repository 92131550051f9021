"""The MSSA and serving workloads at their full sizes against the FP64 oracle,
and the C ABI's refusals that the host can force.

MSSA (MKL/ssa_opt.h:3881-4283): M = 16 channels, N = 10000, L = 2500,
k = 64, p = 10, q = 2, components 0..8 of every channel.
  plain FP64: singular values 1e-10 relative, reconstructions 1e-9 of RMS
  mixed:      singular values 1e-5 relative; the M x N reconstruction and
              each channel 1e-5 relative (2-norm), 1e-4 of the channel's RMS
              pointwise
Serving: requests with N log-uniform in [500, 50000], L = N // 4,
k ~ U{10..50}, group {0..floor(2k/5)-1}, R-forecast h = 50
(MKL/ssa_opt.h:2843-2867); a sample of the served responses, reconstruction
and forecast, against the oracle with each request's own Philox key.
  plain FP64: reconstruction and forecast 1e-9 of RMS(x)
  mixed:      reconstruction 1e-5 relative (2-norm) and 1e-4 of RMS(x)
              pointwise; forecast 1e-4 of RMS(x)
"""
import ctypes

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


@pytest.fixture(scope="module")
def mssa_oracle():
    X = workloads.mssa_channels(M=16, N=10000, seed=5)
    L, k, p = 2500, 64, 10
    K = 10000 - L + 1
    # the device's default Philox key (42) and first call
    o = O.OracleMSSA(X, L).decompose(k, p, 2, O.philox_gaussian(42, 0, K, k + p))
    return X, o


@pytest.mark.parametrize("prec", [0, 1])
def test_mssa_full_size(mssa_oracle, prec):
    X, o = mssa_oracle
    S = _ssa()
    m = S.MSSA(X, 2500, q=2, precision=prec)
    m.decompose(64, oversampling=10)
    s = o.factors()[2]
    dev_s = np.array([m.variance_explained(i, i) for i in range(64)])
    assert np.allclose(dev_s, s ** 2 / np.sum(s ** 2), rtol=1e-9 if prec == 0 else 1e-5)
    g = list(range(9))
    worst = 0.0
    dev_all, orc_all = [], []
    for ch in range(16):
        r = rms(X[ch])
        a, b = m.reconstruct(ch, g), o.reconstruct(ch, g)
        dev_all.append(a)
        orc_all.append(b)
        if prec == 0:
            assert np.abs(a - b).max() / r < 1e-9, ch
            for c in range(9):
                assert np.abs(m.reconstruct(ch, [c]) - o.reconstruct(ch, [c])).max() / r < 1e-9, (ch, c)
        else:
            assert rel2(a, b) < 1e-5, ch
            assert np.abs(a - b).max() / r < 1e-4, ch
        worst = max(worst, rel2(a, b))
    joint = rel2(np.array(dev_all), np.array(orc_all))
    assert joint < (1e-12 if prec == 0 else 1e-5)
    allc = m.reconstruct_all(g)
    for ch in range(16):
        assert np.array_equal(allc[ch], dev_all[ch])
    print(f"MSSA M=16 N=10000 prec={prec}: M x N reconstruction rel2 {joint:.2e}, worst channel {worst:.2e}")


def _serve(prec, reqs):
    import serve
    srv = serve.Server(max_wave=512, inflight=2, precision=prec, hbm_budget=16 * 2 ** 30)
    return srv.run(reqs)


@pytest.mark.parametrize("prec", [1, 0])
def test_serving_full_distribution(prec):
    """Every response of a 256-request backlog (N up to 50000) is served
    (status 0 or 2, the verticality failure); a sample including the longest
    and shortest requests matches the oracle."""
    import serve
    reqs = serve.synthetic_requests(256, rate=np.inf, seed=6)
    done = _serve(prec, reqs)
    assert len(done) == 256
    by_id = {r.rid: r for r in done}
    Ns = np.array([len(r.x) for r in reqs])
    assert Ns.max() > 40000
    pick = sorted(set(np.argsort(Ns)[-4:].tolist() + np.argsort(Ns)[:2].tolist() + list(range(0, 256, 64))))
    worst_r = worst_f = worst_p = 0.0
    for i in pick:
        r = by_id[i]
        assert r.status in (0, 2), (i, r.status)
        x = np.asarray(r.x, dtype=np.float64)
        N, L, k, g = len(x), r.L, r.k, r.g
        K = N - L + 1
        b = min(k + 10, L, K)
        orc = O.OracleSSA(x, L).decompose(k, 10, 2, O.philox_gaussian(serve.request_key(i, 6), 0, K, b))
        grp = list(range(g))
        rec = orc.reconstruct(grp)
        rc, _, _ = orc.lrf(grp)
        assert (r.status == 2) == (rc != 0), (i, r.status, rc)
        s = rms(x)
        worst_r = max(worst_r, rel2(r.out[:N], rec))
        worst_p = max(worst_p, float(np.abs(r.out[:N] - rec).max() / s))
        if r.status == 0:
            worst_f = max(worst_f, float(np.abs(r.out[N:] - orc.forecast(grp, 50)).max() / s))
    print(f"serving prec={prec}: {len(pick)} sampled responses, N up to {Ns.max()}: reconstruction rel2 "
          f"{worst_r:.2e}, pointwise/RMS {worst_p:.2e}, forecast/RMS {worst_f:.2e}")
    if prec == 0:
        assert worst_p < 1e-9 and worst_f < 1e-9
    else:
        assert worst_r < 1e-5 and worst_p < 1e-4 and worst_f < 1e-4


def test_state_machine_refusals():
    """The reference's refusals (MKL/ssa_opt.h:1697-1698, :1414-1543): a
    randomized decomposition needs prepare and kp <= prepared_kp; every
    failure returns -1 and leaves the context usable."""
    S = _ssa()
    x = workloads.single(1000, seed=1)
    s = S.SSA(x, 250)
    lib, ctx = s.lib, s.ref
    assert lib.ssa_opt_decompose_randomized(ctx, 10, 10) == -1          # not prepared
    assert lib.ssa_opt_prepare(ctx, 10, 10) == 0
    assert lib.ssa_opt_decompose_randomized(ctx, 20, 10) == -1          # kp 30 > prepared 20
    assert lib.ssa_opt_decompose_randomized(ctx, 10, 10) == 0
    out = np.zeros(1000)
    g = np.array([0], dtype=np.int32)
    gp = g.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    op = out.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.ssa_opt_reconstruct(ctx, gp, 0, op) == -1               # n_group < 1
    assert lib.ssa_opt_reconstruct(ctx, gp, 1, op) == 0
    assert lib.ssa_opt_update_signal(ctx, x.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    assert lib.ssa_opt_reconstruct(ctx, gp, 1, op) == -1               # update_signal clears decomposed


def test_block_width_limit():
    """Block widths above 80 are refused with -1 and the context stays
    usable: prepare and decompose_randomized need k + p <= 80, decompose
    k + 8 <= 80; decompose_block clamps its block to 80 and needs k <= 80."""
    S = _ssa()
    x = workloads.single(2000, seed=1)
    s = S.SSA(x, 500)
    lib, ctx = s.lib, s.ref
    assert lib.ssa_opt_prepare(ctx, 75, 10) == -1
    assert lib.ssa_opt_decompose(ctx, 80, 20) == -1
    assert lib.ssa_opt_decompose_block(ctx, 81, 16, 20) == -1
    assert lib.ssa_opt_decompose_block(ctx, 76, 16, 20) == 0            # block 84 clamped to 80
    assert lib.ssa_opt_prepare(ctx, 70, 10) == 0
    assert lib.ssa_opt_decompose_randomized(ctx, 70, 10) == 0
    with pytest.raises(RuntimeError):
        s.decompose(80, oversampling=8)
