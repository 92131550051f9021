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
