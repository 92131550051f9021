"""Request/response frames of the serving loop, and the stream-serving loop's
host logic (CPU: the device step is replaced by a stand-in that answers each
request with a function of its own series)."""
import io
import struct

import numpy as np
import pytest

import wire


class Trickle(io.RawIOBase):
    """A stream that hands out at most a few bytes per read, as a socket may."""

    def __init__(self, data, seed=0):
        self.data, self.pos = data, 0
        self.rng = np.random.default_rng(seed)

    def readable(self):
        return True

    def read(self, n=-1):
        m = min(n, int(self.rng.integers(1, 8)), len(self.data) - self.pos)
        out = self.data[self.pos:self.pos + m]
        self.pos += m
        return out


def test_header_sizes_and_layout():
    frame = wire.encode_request(0x0102030405060708, np.arange(5, dtype=np.float32), 3, 2, 1, 7)
    assert len(frame) == 32 + 4 * 5
    assert frame[:4] == b"SSAQ"
    assert struct.unpack_from("<HHQIIHHHH", frame, 4) == (1, 1, 0x0102030405060708, 5, 3, 2, 1, 7, 0)
    assert np.array_equal(np.frombuffer(frame[32:], dtype="<f4"), np.arange(5, dtype=np.float32))
    resp = wire.encode_response(9, 0, np.array([1.5, -2.0]))
    assert len(resp) == 24 + 16 and resp[:4] == b"SSAR"
    assert struct.unpack_from("<HHQiI", resp, 4) == (1, 0, 9, 0, 2)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_request_round_trip_through_a_trickling_stream(seed):
    rng = np.random.default_rng(seed)
    frames, want = b"", []
    for rid in range(6):
        N = int(rng.integers(4, 3000))
        x = rng.standard_normal(N).astype(np.float32)
        L, k = int(rng.integers(2, N)), int(rng.integers(1, 60))
        g, h = int(rng.integers(1, k + 1)), int(rng.integers(0, 100))
        frames += wire.encode_request(rid * 977 + 2**40, x, L, k, g, h)
        want.append((rid * 977 + 2**40, x, L, k, g, h))
    s = Trickle(frames, seed)
    for w in want:
        got = wire.read_request(s)
        assert got[0] == w[0] and got[2:] == w[2:]
        assert got[1].dtype == np.float32 and np.array_equal(got[1], w[1])
    assert wire.read_request(s) is None


def test_response_round_trip():
    out = np.random.default_rng(3).standard_normal(1050)
    data = wire.encode_response(5, 0, out) + wire.encode_response(6, -3) + wire.encode_response(7, -1, out)
    s = Trickle(data, 4)
    rid, st, o = wire.read_response(s)
    assert (rid, st) == (5, 0) and np.array_equal(o, out)
    assert wire.read_response(s)[:2] == (6, -3)
    rid, st, o = wire.read_response(s)
    assert (rid, st, o.size) == (7, -1, 0)          # a failure carries no output
    assert wire.read_response(s) is None


def test_malformed_frames_raise():
    good = wire.encode_request(1, np.ones(10, np.float32), 4, 2, 1, 0)
    with pytest.raises(wire.WireError, match="magic"):
        wire.read_request(io.BytesIO(b"XXXX" + good[4:]))
    with pytest.raises(wire.WireError, match="version"):
        wire.read_request(io.BytesIO(good[:4] + struct.pack("<H", 2) + good[6:]))
    with pytest.raises(wire.WireError, match="sample type"):
        wire.read_request(io.BytesIO(good[:6] + struct.pack("<H", 2) + good[8:]))
    with pytest.raises(wire.WireError, match="inside a frame"):
        wire.read_request(io.BytesIO(good[:20]))
    with pytest.raises(wire.WireError):
        wire.read_request(io.BytesIO(good[:-3]))                  # series cut short
    big = bytearray(good)
    struct.pack_into("<I", big, 16, wire.MAX_N + 1)
    with pytest.raises(wire.WireError, match="above"):
        wire.read_request(io.BytesIO(bytes(big)))
    with pytest.raises(wire.WireError, match="magic"):
        wire.read_response(io.BytesIO(b"SSAQ" + bytes(20)))


def test_field_ranges():
    assert wire.check_fields(1000, 250, 50, 20, 50) is None
    assert wire.check_fields(4, 2, 1, 1, 0) is None
    assert "N" in wire.check_fields(3, 2, 1, 1, 0)
    assert "L" in wire.check_fields(100, 1, 1, 1, 0)
    assert "L" in wire.check_fields(100, 100, 1, 1, 0)
    assert "k" in wire.check_fields(100, 25, 0, 1, 0)
    assert wire.check_fields(100, 25, 70, 1, 0, p=10) is None
    assert "k" in wire.check_fields(100, 25, 71, 1, 0, p=10)
    assert "g" in wire.check_fields(100, 25, 5, 6, 0)
    assert "g" in wire.check_fields(100, 25, 5, 0, 0)


class StandIn:
    """A Server whose device step is replaced: each request's output is its
    series (as float64) followed by h copies of its last sample; a series
    whose first sample is negative gets device status 2."""

    @staticmethod
    def make(**kw):
        import serve

        class S(serve.Server):
            def _launch(self, reqs):
                self.waves.append((None, reqs))
                self.wave_sizes.append(len(reqs))

            def _retire(self, block):
                _, reqs = self.waves.pop(0)
                for r in reqs:
                    r.out = np.concatenate([r.x.astype(np.float64), np.full(self.h, float(r.x[-1]))])
                    r.status = 2 if r.x[0] < 0 else 0
                    import time
                    r.done = time.perf_counter()
                return True

            def footprint(self, r):
                return 8 * len(r.x)

        return S(**kw)


def test_serve_stream_answers_every_frame():
    rng = np.random.default_rng(7)
    frames, sent = b"", {}
    for rid in range(25):
        N = int(rng.integers(10, 400))
        x = (rng.standard_normal(N) + 5).astype(np.float32)
        L, k, g, h = N // 4, 12, 4, 5
        if rid == 3:
            L = N                  # out of range: rejected
        if rid == 4:
            h = 6                  # not the server's horizon: rejected
        if rid == 5:
            x[0] = -1.0            # device status 2: failed
        if rid == 6:
            N = 5000               # above the 1 KiB budget's whole: refused
            x = np.ones(N, np.float32)
            L = 100
        frames += wire.encode_request(rid, x, L, k, g, h)
        sent[rid] = x
    srv = StandIn.make(h=5, max_wave=4, inflight=2, hbm_budget=8 * 1024)
    out = io.BytesIO()
    done, err = srv.serve_stream(Trickle(frames, 8), out)
    assert err is None and len(done) == 25
    out.seek(0)
    got = {}
    while True:
        r = wire.read_response(out)
        if r is None:
            break
        got[r[0]] = r[1:]
    assert sorted(got) == list(range(25))
    assert got[3][0] == got[4][0] == wire.STATUS_REJECTED
    assert got[5][0] == wire.STATUS_FAILED and got[5][1].size == 0
    assert got[6][0] == wire.STATUS_REFUSED
    for rid, x in sent.items():
        if rid in (3, 4, 5, 6):
            continue
        st, o = got[rid]
        assert st == 0 and np.array_equal(o[:len(x)], x.astype(np.float64)) and o.size == len(x) + 5
    assert max(srv.wave_sizes) <= 4


def test_serve_stream_answers_a_wave_the_device_cannot_take():
    """A wave whose launch fails is answered with status -1, and the
    requests after it are still served."""
    srv = StandIn.make(h=5, max_wave=1, inflight=1)
    launch = srv._launch

    def flaky(reqs):
        if reqs[0].rid == 1:
            raise RuntimeError("batch creation failed")
        launch(reqs)

    srv._launch = flaky
    frames = b"".join(wire.encode_request(i, np.ones(40, np.float32) * (i + 1), 10, 3, 1, 5) for i in range(3))
    out = io.BytesIO()
    done, err = srv.serve_stream(io.BytesIO(frames), out)
    assert err is None and len(done) == 3
    out.seek(0)
    got = {}
    for _ in range(3):
        rid, st, o = wire.read_response(out)
        got[rid] = (st, o.size)
    assert got == {0: (0, 45), 1: (wire.STATUS_FAILED, 0), 2: (0, 45)}


def test_serve_stream_stops_at_a_broken_frame():
    good = b"".join(wire.encode_request(i, np.ones(50, np.float32) * (i + 1), 10, 3, 1, 5) for i in range(3))
    srv = StandIn.make(h=5, max_wave=8, inflight=1)
    out = io.BytesIO()
    done, err = srv.serve_stream(io.BytesIO(good + b"JUNK" + bytes(40)), out)
    assert isinstance(err, wire.WireError) and len(done) == 3
    out.seek(0)
    rids = sorted(wire.read_response(out)[0] for _ in range(3))
    assert rids == [0, 1, 2] and wire.read_response(out) is None
