"""Serving loop: a request's result does not depend on the wave it is served
in (alone vs in the mix), and every request of an open-loop stream completes
with its group reconstruction and forecast."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _serve():
    import serve
    return serve


def test_alone_vs_in_the_mix_bit_identical():
    S = _serve()
    reqs = S.synthetic_requests(10, rate=np.inf, n_lo=500, n_hi=6000, seed=31)
    mixed = S.Server(max_wave=10, inflight=1).run(reqs)
    by_id = {r.rid: r for r in mixed}
    for r in S.synthetic_requests(10, rate=np.inf, n_lo=500, n_hi=6000, seed=31):
        alone = S.Server(max_wave=1, inflight=1).run([r])[0]
        m = by_id[r.rid]
        assert m.out.shape == alone.out.shape == (len(r.x) + r.h,)
        assert np.array_equal(m.out, alone.out), r.rid
        assert m.status == alone.status


def test_branch_count_and_concurrent_waves_bit_identical():
    """A wave split over 1, 2 or 4 graph branches, and the same requests served
    by four one-branch waves in flight at once (their kernels overlap on the
    device), give the same bits: every per-series computation is independent
    of the engine and stream it runs on."""
    S = _serve()
    import ssa_amd
    reqs = S.synthetic_requests(24, rate=np.inf, n_lo=500, n_hi=20000, seed=33)
    xs, Ls, ks = [r.x for r in reqs], [r.L for r in reqs], [r.k for r in reqs]
    gs, keys = [r.g for r in reqs], [S.request_key(r.rid) for r in reqs]
    ref = None
    for br in (1, 2, 4):
        b = ssa_amd.Batch(xs, Ls, ks, p=10, q=2, precision=1, groups=gs, h=50, keys=keys, branches=br)
        b.set_wide_min(1)
        b.run(use_graph=False)
        b.sync()
        outs, st = b.outputs()
        if ref is None:
            ref = outs
        else:
            for i in range(len(reqs)):
                assert np.array_equal(outs[i], ref[i]), (br, i)
        del b
    served = {r.rid: r for r in S.Server(max_wave=6, inflight=4).run(reqs)}
    for i, r in enumerate(reqs):
        assert np.array_equal(served[r.rid].out, ref[i]), r.rid


def test_alone_vs_a_wave_of_700_bit_identical():
    """A wave of 700 short series launches its Gram, apply and test-matrix
    kernels one workgroup per row tile; a request served alone deals each
    tile over several workgroups.  Every output row and every Gram block is
    formed in the same order either way, so the results are bit-identical."""
    S = _serve()
    reqs = S.synthetic_requests(700, rate=np.inf, n_lo=500, n_hi=1500, seed=35)
    big = {r.rid: r for r in S.Server(max_wave=700, inflight=1).run(reqs)}
    for r in S.synthetic_requests(700, rate=np.inf, n_lo=500, n_hi=1500, seed=35)[::97]:
        alone = S.Server(max_wave=1, inflight=1).run([r])[0]
        assert np.array_equal(big[r.rid].out, alone.out), r.rid
        assert big[r.rid].status == alone.status


@pytest.mark.parametrize("prec,count,n_lo,n_hi", [(1, 48, 25000, 50000), (0, 300, 3000, 8000)])
def test_alone_vs_a_wave_on_the_wide_geometry_bit_identical(prec, count, n_lo, n_hi):
    """Waves whose block-matvec launches hold thousands of items: FP32 series
    at n = 8192 chunked on both sides, FP64 series at n = 1024..4096 whose K
    dimension is split.  With the launch-size rule a batch would run these
    launches on the wide transform geometry and a request alone on the narrow
    one, which rounds differently; the server pins the wide geometry, so the
    results are bit-identical."""
    S = _serve()
    reqs = S.synthetic_requests(count, rate=np.inf, n_lo=n_lo, n_hi=n_hi, seed=37)
    big = {r.rid: r for r in S.Server(max_wave=count, inflight=1, precision=prec).run(reqs)}
    for r in S.synthetic_requests(count, rate=np.inf, n_lo=n_lo, n_hi=n_hi, seed=37)[::count // 3]:
        alone = S.Server(max_wave=1, inflight=1, precision=prec).run([r])[0]
        assert np.array_equal(big[r.rid].out, alone.out), (r.rid, len(r.x))
        assert big[r.rid].status == alone.status


def test_open_loop_stream_completes_with_forecasts():
    S = _serve()
    reqs = S.synthetic_requests(40, rate=400.0, n_lo=500, n_hi=8000, seed=32)
    done = S.Server(max_wave=16, inflight=2).run(reqs)
    assert len(done) == 40
    for r in done:
        assert r.done >= r.arrival
        assert np.all(np.isfinite(r.out))
        # the forecast continues the reconstruction at the series' scale
        scale = np.abs(r.out[:len(r.x)]).max() + 1e-12
        assert np.abs(r.out[len(r.x):]).max() < 10 * scale
    st = S.latency_stats(done)
    assert st["requests"] == 40 and st["p99_ms"] >= st["p50_ms"] > 0


def test_rebind_matches_fresh_batch():
    """A batch re-targeted to new series (reusing its device buffers) gives
    bit-identical results to a batch created for those series."""
    import ssa_amd as S
    import workloads
    xa, Na, La, ka = workloads.ragged_batch(count=7, n_lo=500, n_hi=9000, seed=51, dtype=np.float32)
    xb, Nb, Lb, kb = workloads.ragged_batch(count=5, n_lo=800, n_hi=12000, seed=52, dtype=np.float32)
    keys = [1000 + i for i in range(5)]
    gb = [max(1, int(k) // 2) for k in kb]
    a = S.Batch(xa, La, ka, p=10, q=2, precision=1, seed=1, groups=[2] * 7, h=20)
    a.run(use_graph=False)
    a.sync()
    a.rebind(xb, Lb, kb, groups=gb, keys=keys)
    a.run(use_graph=False)
    a.sync()
    fresh = S.Batch(xb, Lb, kb, p=10, q=2, precision=1, seed=1, groups=gb, h=20, keys=keys)
    fresh.run(use_graph=False)
    fresh.sync()
    oa, sa = a.outputs()
    of, sf = fresh.outputs()
    for i in range(5):
        assert oa[i].shape == (int(Nb[i]) + 20,)
        assert np.array_equal(oa[i], of[i]), i
        assert np.array_equal(a.sigma(i), fresh.sigma(i)), i
    assert np.array_equal(sa, sf)


def test_series_bytes_match_the_batch():
    """ssa_amd.series_bytes (host arithmetic) is exactly what a batch's arenas
    use, series by series, including the spectrum scratch of series whose
    block matvec takes the staged form (chunked on both sides, N up to 50000)."""
    import ssa_amd as S
    import workloads
    xs, Ns, Ls, ks = workloads.ragged_batch(count=10, n_lo=500, n_hi=20000, seed=61, dtype=np.float32)
    xl, Nl, Ll, kl = workloads.ragged_batch(count=4, n_lo=30000, n_hi=50000, seed=62, dtype=np.float32)
    xs, Ls, ks = list(xs) + list(xl), np.concatenate([Ls, Ll]), np.concatenate([ks, kl])
    for prec in (0, 1):
        b = S.Batch(xs, Ls, ks, p=10, q=2, precision=prec, seed=1, h=50)
        staged = [b.plan(i) for i in range(len(xs))]
        assert any(pl["AL"] > 1 and pl["SK"] > 1 for pl in staged), prec
        est = sum(S.series_bytes(len(x), int(L), int(k), 10, 2, prec, 50) for x, L, k in zip(xs, Ls, ks))
        assert b.bytes() == est, prec


def test_budgeted_serving_matches_unbudgeted():
    """Waves cut to a device-memory budget give every request the result it
    gets in one unbounded wave (bit for bit), and stay within the share."""
    S = _serve()
    reqs = S.synthetic_requests(16, rate=np.inf, n_lo=500, n_hi=8000, seed=33)
    free = {r.rid: r for r in S.Server(max_wave=16, inflight=1).run(reqs)}
    sv = S.Server(max_wave=16, inflight=2)
    total = sum(sv.footprint(r) for r in S.synthetic_requests(16, rate=np.inf, n_lo=500, n_hi=8000, seed=33))
    sv = S.Server(max_wave=16, inflight=2, hbm_budget=total / 2)
    capped = sv.run(S.synthetic_requests(16, rate=np.inf, n_lo=500, n_hi=8000, seed=33))
    assert len(sv.wave_sizes) >= 3 and sum(sv.wave_sizes) == 16
    for wb, ws in zip(sv.wave_bytes, sv.wave_sizes):
        assert wb <= total / 4 or ws == 1
    for r in capped:
        assert r.status == free[r.rid].status == 0
        assert np.array_equal(r.out, free[r.rid].out), r.rid


def test_stream_frames_match_in_process_serving():
    """Requests sent as wire frames over a socket pair are answered with the
    outputs in-process serving gives the same requests (bit-identical), and a
    frame with an out-of-range field is answered with status -4."""
    import socket
    import threading

    import wire
    S = _serve()
    ref = {r.rid: r for r in S.Server(max_wave=12, inflight=1).run(
        S.synthetic_requests(12, rate=np.inf, n_lo=500, n_hi=6000, seed=33))}
    a, b = socket.socketpair()
    got = {}

    def client():
        with a.makefile("wb") as w:
            for r in S.synthetic_requests(12, rate=np.inf, n_lo=500, n_hi=6000, seed=33):
                w.write(wire.encode_request(r.rid, r.x, r.L, r.k, r.g, r.h))
            w.write(wire.encode_request(99, np.ones(100, np.float32), 100, 5, 1, 50))   # L = N: rejected
        a.shutdown(socket.SHUT_WR)
        with a.makefile("rb") as rd:
            while True:
                f = wire.read_response(rd)
                if f is None:
                    break
                got[f[0]] = f[1:]

    th = threading.Thread(target=client)
    th.start()
    with b.makefile("rb") as rf, b.makefile("wb") as wf:
        done, err = S.Server(max_wave=4, inflight=2).serve_stream(rf, wf)
    b.shutdown(socket.SHUT_WR)
    th.join(timeout=60)
    a.close()
    b.close()
    assert err is None and len(done) == 13
    assert got[99][0] == wire.STATUS_REJECTED
    for rid, r in ref.items():
        st, out = got[rid]
        assert st == (0 if r.status == 0 else wire.STATUS_FAILED)
        if st == 0:
            assert np.array_equal(out, r.out), rid
