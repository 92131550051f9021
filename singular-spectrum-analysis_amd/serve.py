"""Continuous-batching SSA serving loop on one GPU (no counterpart in the
reference, which serves one context per call).

A request is one series x with its window L, components k, a reconstruction
group {0..g-1} and a forecast horizon h: decompose (randomized, q power
iterations) + reconstruct the group + R-forecast h steps (`ssa_opt_forecast`,
MKL/ssa_opt.h:2843-2867, on the group's reconstruction).

The server gathers the requests that have arrived into a wave (at most
`max_wave` series, or with `min_wave` an equal share of the waiting queue
over the in-flight slots, and, with `hbm_budget`, at most
hbm_budget / inflight device bytes by `ssa_amd.series_bytes`), runs the wave as one batched step on
the device (one `ssa_amd_batch`, mixed or plain form) and hands every request
its output.  A request larger than a slot's share runs alone; one larger than
the whole budget is refused (status -3, no output).
Up to `inflight` waves are outstanding (default 4): while they run on the GPU
the next is assembled and uploaded.  Each wave's batch runs on one stream (one
graph branch), so the in-flight waves' streams do not outnumber the device's
4 hardware queues and their kernels overlap.  The server keeps one batch per in-flight slot and
re-targets it to each new wave (`ssa_amd_batch_rebind`), so once the slots
have grown to the largest wave no device memory is allocated or freed.  Each
request carries its own Philox key (its id), so its test matrix, and hence its
result, does not depend on the wave it lands in.

`Server.run` serves an in-process list of requests; `Server.serve_stream`
serves request frames of the binary format in `wire.py` read from a stream
(a socket, a pipe) and answers each with a response frame.
"""
import heapq
import time
from dataclasses import dataclass, field

import numpy as np

import ssa_amd


@dataclass(order=True)
class Request:
    arrival: float
    rid: int
    x: np.ndarray = field(compare=False)
    L: int = field(compare=False)
    k: int = field(compare=False)
    g: int = field(compare=False)
    h: int = field(compare=False)
    done: float = field(default=-1.0, compare=False)
    out: np.ndarray = field(default=None, compare=False)
    status: int = field(default=0, compare=False)
    nbytes: int = field(default=0, compare=False)


def request_key(rid, seed=6):
    return (seed * 0x100000001B3 + 0x9E3779B97F4A7C15 * (rid + 1)) & ((1 << 64) - 1)


class Server:
    """Waves of requests through `ssa_amd.Batch` (all requests of a wave share
    h, p, q and the precision)."""

    def __init__(self, p=10, q=2, h=50, precision=1, max_wave=512, inflight=4, seed=6, hbm_budget=None,
                 branches=1, min_wave=None):
        self.p, self.q, self.h, self.precision = p, q, h, precision
        self.max_wave, self.inflight, self.seed = max_wave, inflight, seed
        self.min_wave = min_wave
        self.branches = branches
        self.hbm_budget = hbm_budget
        self.waves = []          # (batch, requests) in flight
        self.idle = []           # batches ready for the next wave
        self.wave_sizes = []
        self.wave_bytes = []     # device bytes of each wave's series

    def footprint(self, r):
        if not r.nbytes:
            r.nbytes = ssa_amd.series_bytes(len(r.x), r.L, r.k, self.p, self.q, self.precision, self.h)
        return r.nbytes

    def wave_cap(self, n_waiting):
        """Requests one wave may take: max_wave, or with min_wave set an equal
        share of the queue over the in-flight slots, at least min_wave and at
        most max_wave (a wave then leaves part of a burst to the next free
        slot, while a backlog still fills waves of max_wave).  On the serving
        distribution at 20k and 28.8k requests/s the share rule (min_wave 32,
        64, 128) moved p50 and p99 by less than the run-to-run spread, so it
        is off by default."""
        if self.min_wave is None:
            return self.max_wave
        return min(self.max_wave, max(self.min_wave, -(-n_waiting // self.inflight)))

    def next_wave(self, waiting, cap=None):
        """The longest prefix of `waiting` (arrival order) that fits a wave of
        at most `cap` requests (default wave_cap): (wave, rest, refused)."""
        if cap is None:
            cap = self.wave_cap(len(waiting))
        if self.hbm_budget is None:
            return waiting[:cap], waiting[cap:], []
        share = self.hbm_budget / self.inflight
        wave, used, refused = [], 0, []
        i = 0
        while i < len(waiting) and len(wave) < cap:
            r = waiting[i]
            need = self.footprint(r)
            if need > self.hbm_budget:
                refused.append(r)
            elif not wave or used + need <= share:
                wave.append(r)
                used += need
                if used > share:       # a request above the share runs alone
                    i += 1
                    break
            else:
                break
            i += 1
        return wave, waiting[i:], refused

    def warm(self, requests):
        """Grow every in-flight slot to hold the first wave `requests` would
        form (device arenas, pinned staging), outside any timed run."""
        wave, _, _ = self.next_wave(list(requests), cap=self.max_wave)
        for _ in range(self.inflight):
            self._launch(list(wave))
        while self.waves:
            self._retire(block=True)
        self.wave_sizes.clear()
        self.wave_bytes.clear()

    def _launch(self, reqs):
        xs, Ls, ks = [r.x for r in reqs], [r.L for r in reqs], [r.k for r in reqs]
        gs, keys = [r.g for r in reqs], [request_key(r.rid, self.seed) for r in reqs]
        if self.idle:
            b = self.idle.pop()
            b.rebind(xs, Ls, ks, groups=gs, keys=keys)
        else:
            b = ssa_amd.Batch(xs, Ls, ks, p=self.p, q=self.q, precision=self.precision, groups=gs, h=self.h,
                              keys=keys, branches=self.branches)
            # one transform geometry for every wave size: a request's result
            # then does not depend on the wave it lands in
            b.set_wide_min(1)
        b.run(use_graph=False)
        self.waves.append((b, reqs))
        self.wave_sizes.append(len(reqs))
        self.wave_bytes.append(sum(self.footprint(r) for r in reqs) if self.hbm_budget is not None else 0)

    def _retire(self, block, j=0):
        """Hand out the outputs of in-flight wave j once it has finished
        (block: wait for it); True if it was retired."""
        b, reqs = self.waves[j]
        if not block and not b.query():
            return False
        b.sync()
        outs, status = b.outputs()
        now = time.perf_counter()
        for i, r in enumerate(reqs):
            r.out = outs[i]
            r.status = int(status[i])
            r.done = now
        self.waves.pop(j)
        self.idle.append(b)
        return True

    def _retire_finished(self):
        """Retire every in-flight wave that has finished, in any order (a
        small wave can finish before an older large one); how many."""
        n, j = 0, 0
        while j < len(self.waves):
            if self._retire(block=False, j=j):
                n += 1
            else:
                j += 1
        return n

    def run(self, requests, t0=None):
        """Serve `requests` (arrival times relative to t0, seconds) in real
        time; returns them with completion times set."""
        pending = list(requests)
        heapq.heapify(pending)
        t0 = time.perf_counter() if t0 is None else t0
        waiting = []
        finished = []
        while pending or waiting or self.waves:
            now = time.perf_counter() - t0
            while pending and pending[0].arrival <= now:
                waiting.append(heapq.heappop(pending))
            if waiting and len(self.waves) < self.inflight:
                wave, waiting, refused = self.next_wave(waiting)
                for r in refused:
                    r.status, r.done = -3, time.perf_counter()
                finished.extend(refused)
                if wave:
                    self._launch(wave)
                    finished.extend(wave)
                continue
            if self.waves and self._retire_finished():
                continue
            # nothing ready: sleep until the next arrival or poll the waves
            nap = 2e-4
            if pending and not self.waves:
                nap = max(0.0, min(0.01, pending[0].arrival - now))
            time.sleep(nap)
        for r in finished:
            r.done -= t0
        return finished


    def serve_stream(self, rfile, wfile):
        """Serve request frames read from the binary stream `rfile` until it
        ends, writing one response frame per request to `wfile` as its wave
        retires (completion order; the request id pairs them).  Frames are
        read on a thread of their own, so requests keep arriving while waves
        run.  Returns (served requests, WireError or None): a frame that cannot
        be parsed ends the reading, and the requests read before it are still
        answered."""
        import queue
        import threading

        import wire
        inbox = queue.Queue()
        state = {"error": None}
        t0 = time.perf_counter()

        def reader():
            try:
                while True:
                    f = wire.read_request(rfile)
                    if f is None:
                        break
                    rid, x, L, k, g, h = f
                    why = wire.check_fields(len(x), L, k, g, h, self.p)
                    if why is None and h != self.h:
                        why = f"h = {h}, this server forecasts {self.h}"
                    r = Request(time.perf_counter() - t0, rid, x, L, k, g, h)
                    if why is not None:
                        r.status = wire.STATUS_REJECTED
                    inbox.put(r)
            except wire.WireError as exc:
                state["error"] = exc
            inbox.put(None)

        th = threading.Thread(target=reader, daemon=True)
        th.start()
        waiting, finished, reading = [], [], True

        def answer(rs):
            for r in rs:
                # a device status word other than 0 (2: verticality >= 1, no forecast) is a failure
                st = r.status if r.status in (wire.STATUS_OK, wire.STATUS_REFUSED, wire.STATUS_REJECTED) \
                    else wire.STATUS_FAILED
                wfile.write(wire.encode_response(r.rid, st, r.out))
            wfile.flush()

        while reading or waiting or self.waves:
            try:
                while True:
                    r = inbox.get(block=not (waiting or self.waves), timeout=None)
                    if r is None:
                        reading = False
                        break
                    if r.status == wire.STATUS_REJECTED:
                        r.done = time.perf_counter() - t0
                        answer([r])
                        finished.append(r)
                    else:
                        waiting.append(r)
                    if not reading or inbox.empty():
                        break
            except queue.Empty:
                pass
            if waiting and len(self.waves) < self.inflight:
                wave, waiting, refused = self.next_wave(waiting)
                for r in refused:
                    r.status, r.done = wire.STATUS_REFUSED, time.perf_counter() - t0
                answer(refused)
                finished.extend(refused)
                if wave:
                    try:
                        self._launch(wave)
                    except RuntimeError:
                        # the device could not take the wave (allocation or set-up failed):
                        # its requests are answered as failed and serving goes on
                        for r in wave:
                            r.status, r.done = wire.STATUS_FAILED, time.perf_counter() - t0
                        answer(wave)
                        finished.extend(wave)
                continue
            if self.waves:
                reqs = self.waves[0][1]
                # block on the oldest wave when no slot is free or nothing more can arrive;
                # otherwise poll, so that a new arrival takes the free slot at once
                if self._retire(block=len(self.waves) >= self.inflight or not reading):
                    for r in reqs:
                        r.done -= t0
                    answer(reqs)
                    finished.extend(reqs)
                    continue
                time.sleep(2e-4)
        th.join()
        return finished, state["error"]


def synthetic_requests(count, rate, n_lo=500, n_hi=50000, k_lo=10, k_hi=50, h=50, seed=6, t_start=0.0):
    """Open-loop Poisson arrivals at `rate` per second: N log-uniform in
    [n_lo, n_hi], L = N // 4, k ~ U{k_lo..k_hi}, group {0..floor(2k/5)-1},
    signals of the benchmark family (FP32).  rate = inf: all at t_start."""
    import workloads
    xs, Ns, Ls, ks = workloads.ragged_batch(count=count, n_lo=n_lo, n_hi=n_hi, k_lo=k_lo, k_hi=k_hi, seed=seed,
                                            dtype=np.float32)
    rng = np.random.default_rng(seed + 1000)
    if np.isinf(rate):
        arrivals = np.full(count, t_start)
    else:
        arrivals = t_start + np.cumsum(rng.exponential(1.0 / rate, size=count))
    return [Request(float(arrivals[i]), i, xs[i], int(Ls[i]), int(ks[i]), max(1, (2 * int(ks[i])) // 5), h)
            for i in range(count)]


def latency_stats(reqs):
    """Latency percentiles over the served requests (refused ones excluded)."""
    reqs = [r for r in reqs if r.status != -3]
    lat = np.array([r.done - r.arrival for r in reqs])
    return {"p50_ms": round(1e3 * float(np.percentile(lat, 50)), 3),
            "p99_ms": round(1e3 * float(np.percentile(lat, 99)), 3),
            "mean_ms": round(1e3 * float(lat.mean()), 3), "requests": len(reqs)}
