"""Admission against a device-memory budget (host logic, no device call):
ssa_amd.series_bytes follows the engine layout's growth, and the serving
loop's waves respect the per-slot share of the budget."""
import numpy as np
import pytest


def _S():
    import ssa_amd
    return ssa_amd


def test_series_bytes_scaling():
    S = _S()
    small = S.series_bytes(1000, 250, 20)
    assert small > 0
    assert S.series_bytes(2000, 500, 20) > small
    assert S.series_bytes(1000, 250, 40) > small
    big_mixed = S.series_bytes(20000, 5000, 50, precision=1)
    big_plain = S.series_bytes(20000, 5000, 50, precision=0)
    assert 1.8 < big_plain / big_mixed < 2.1      # FP64 panels and spectra: twice the bytes
    assert S.series_bytes(1000, 250, 20, h=500) > small
    for bad in ((3, 2, 1), (100, 100, 5), (100, 20, 0)):
        with pytest.raises(ValueError):
            S.series_bytes(*bad)


def test_waves_respect_the_budget():
    import serve
    reqs = serve.synthetic_requests(60, rate=np.inf, n_lo=500, n_hi=20000, seed=8)
    sv = serve.Server(max_wave=64, inflight=2, hbm_budget=200e6)
    share = 100e6
    waiting, waves = list(reqs), []
    while waiting:
        wave, waiting, refused = sv.next_wave(waiting)
        assert not refused and wave
        waves.append(wave)
    assert sum(len(w) for w in waves) == 60
    for w in waves:
        used = sum(sv.footprint(r) for r in w)
        assert used <= share or len(w) == 1
    # arrival order is kept
    assert [r.rid for w in waves for r in w] == [r.rid for r in reqs]
    # a request beyond the whole budget is refused, one beyond the share runs alone
    tiny = serve.Server(max_wave=64, inflight=2, hbm_budget=1.5 * sv.footprint(reqs[0]))
    big = max(reqs, key=tiny.footprint)
    wave, rest, refused = tiny.next_wave([big] + [r for r in reqs if r is not big])
    assert (refused and refused[0] is big) or (len(wave) == 1 and wave[0] is big)


def test_wave_cap_shares_the_queue_over_slots():
    """Without min_wave a wave takes up to max_wave requests; with it, an
    equal share of the queue over the in-flight slots, between min_wave and
    max_wave, and a backlog still fills waves of max_wave."""
    import serve
    plain = serve.Server(max_wave=512, inflight=4)
    assert [plain.wave_cap(n) for n in (1, 100, 5000)] == [512, 512, 512]
    shared = serve.Server(max_wave=512, inflight=4, min_wave=64)
    assert [shared.wave_cap(n) for n in (1, 100, 256, 300, 1000, 5000)] == [64, 64, 64, 75, 250, 512]
    reqs = serve.synthetic_requests(300, rate=np.inf, n_lo=500, n_hi=2000, seed=9)
    wave, rest, refused = shared.next_wave(list(reqs))
    assert len(wave) == 75 and len(rest) == 225 and not refused
    assert [r.rid for r in wave + rest] == [r.rid for r in reqs]
    # the warm-up grows every slot to a full wave
    w, _, _ = shared.next_wave(list(reqs), cap=shared.max_wave)
    assert len(w) == 300
