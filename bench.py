"""Throughput of batched SSA on one MI355X.

Workload (the metric's configuration): a ragged batch of 4096 series,
N log-uniform in [500, 20000], L = N // 4, k ~ U{10..50}, oversampling
p = 10, q = 2 power iterations, FP32 inputs.  One step = one SSA of every
series in one HIP-graph launch: segment spectra of the (HBM-resident)
signals, randomized decomposition, reconstruction of the group {0..k-1}.
`value` is SSA/s of the mixed form (FP32 bulk, FP64 small solves); the
plain FP64 form is timed beside it.

Prints one JSON line.  --gpus N > 1 is refused: this path runs on one GPU.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for p in (os.path.join(ROOT, "singular-spectrum-analysis_amd"), os.path.join(ROOT, "oracle")):
    if p not in sys.path:
        sys.path.insert(0, p)

import workloads  # noqa: E402

ssa_amd = None   # loaded after the CPU baseline has forked its workers

HBM_PEAK_GBS = 8000.0
P_OVERSAMPLE = 10
Q_POWER = 2
# the reference's published decompose + reconstruct times (k = 50, p = 8,
# L = N // 4, i9-14900KF, MKL 2025.0.3; its README.md:49-63), ms by N
README_SHAPES_MS = {500: 2.0, 1000: 1.6, 2000: 2.5, 5000: 7.2, 10000: 12.2, 20000: 26.3}


def algorithmic_bytes(N, L, k, b, s):
    """Per-SSA algorithmic HBM bytes: signal read, 6 block-matvec reads and
    writes, 5 orthonormalised panels read and written once, B and Q read for
    the projected SVD and U, U and V written, one reconstruction written."""
    K = N - L + 1
    return s * (N + 6 * b * (L + K) + 2 * b * (3 * L + 2 * K) + b * (L + K) + k * (L + K) + N)


def matvec_bytes(N, L, b, s):
    """One block matvec of one series: the b input and b output columns plus
    the signal spectrum (N/2+1 complex)."""
    K = N - L + 1
    return s * b * (L + K) + 2 * s * (N // 2 + 1)


def timed_steps(batch, steps, warmup):
    for _ in range(warmup):
        batch.run(use_graph=True)
    batch.sync()
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        batch.run(use_graph=True)
        batch.sync()
        lat.append(time.perf_counter() - a)
    total = time.perf_counter() - t0
    return total, np.array(lat)


def step_check(batch):
    """After the timed steps: every series' device status word and whether
    its output is finite (the parity of these outputs with the oracle is
    tests/test_gpu_wide.py's)."""
    outs, status = batch.outputs()
    bad = int(np.count_nonzero(status))
    nonfinite = int(sum(1 for o in outs if not np.all(np.isfinite(o))))
    if bad or nonfinite:
        print(f"[bench] WARNING: {bad} series with a failure status, {nonfinite} with non-finite output",
              file=sys.stderr, flush=True)
    return {"series": len(outs), "status_nonzero": bad, "nonfinite": nonfinite}


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` on this workload, from the rocprofv3
    FETCH_SIZE / WRITE_SIZE passes recorded in profiles/traffic.json (each
    counter scaled by what it reads for a copy kernel of known size), or None."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    try:
        with open(path) as f:
            entry = json.load(f).get(kernel)
    except (OSError, ValueError):
        return None
    if not entry:
        return None
    return float(entry["fetch_kib"]) * 1024 * float(entry["fetch_scale"]) + \
        float(entry["write_kib"]) * 1024 * float(entry["write_scale"])


def measured_valu_issue(kernel):
    """Fraction of the VALU issue rate `kernel` reaches on this workload, from
    the rocprofv3 SQ_INSTS_VALU / GRBM_GUI_ACTIVE pass recorded in
    profiles/traffic.json, or None."""
    try:
        with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
            entry = json.load(f).get(kernel) or {}
    except (OSError, ValueError):
        return None
    return entry.get("valu_issue")


def roofline_of(batch, plans, Ns, Ls, s):
    """Live HIP-event timing of every launch of one eager step; the dominant
    kernel's algorithmic bytes per launch over its average duration."""
    recs = batch.profile()
    agg = {}
    for ph, lg, ms in recs:
        agg.setdefault((ph, lg), []).append(ms)
    total = sum(sum(v) for v in agg.values())
    phase_tot = {}
    for (ph, lg), v in agg.items():
        phase_tot[ph] = phase_tot.get(ph, 0.0) + sum(v)
    # dominant kernel: the k_matvec instantiation of the bucket with the largest
    # share (both directions move the same bytes); the benchmark batch gives
    # each branch's launches of that bucket thousands of items, so they take
    # the wide geometry (k_matvec<T, lg, true>)
    mv = {}
    for (ph, lg), v in agg.items():
        if ph in ("matvec_H", "matvec_HT"):
            mv.setdefault(lg, []).extend(v)
    lg, v = max(mv.items(), key=lambda kv: sum(kv[1]))
    sel = [i for i, p in enumerate(plans) if p["lgn"] == lg]
    # a step applies H or H^T 2(q+1) times to every series of the bucket; the
    # bucket may be split over several launches (parallel branches)
    step_bytes = 2 * (Q_POWER + 1) * float(sum(matvec_bytes(int(Ns[i]), int(Ls[i]), plans[i]["b"], s) for i in sel))
    bytes_launch = step_bytes / len(v)
    avg_ms = float(np.mean(v))
    achieved = step_bytes / (float(np.sum(v)) * 1e-3) / 1e9
    kernel = f"k_matvec<{'float' if s == 4 else 'double'}, {lg}, true>"
    traffic = measured_traffic(kernel)
    return {
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "kernel": kernel,
                     "bytes_per_launch": bytes_launch, "avg_launch_ms": round(avg_ms, 4),
                     "series_in_bucket": len(sel), "launches_timed": len(v),
                     # the kernel is bound by its FFT arithmetic: the VALU issue rate it reaches
                     "valu_issue": measured_valu_issue(kernel)},
        "phase_ms": {k: round(v, 3) for k, v in sorted(phase_tot.items(), key=lambda kv: -kv[1])},
        "profiled_step_ms": round(total, 3),
    }


def _oracle_ssa(args):
    """One SSA of series i by the FP64 oracle: decompose q=2 p=10 with the
    series' Philox test matrix, reconstruct all k components."""
    x, L, k = args
    import oracle as O
    x = np.asarray(x, dtype=np.float64)
    K = len(x) - L + 1
    b = min(k + P_OVERSAMPLE, L, K)
    s = O.OracleSSA(x, L).decompose(k, P_OVERSAMPLE, Q_POWER, O.philox_gaussian(4, 0, K, b))
    s.reconstruct(list(range(k)))
    return len(x)


def _ref_ssa(args):
    """One SSA of series i by the reference's own dependency-free build
    (MKL/ssa_opt_ref.h compiled from source): its decompose_randomized (q = 0,
    the reference has no power iterations; p = 10, its xorshift test matrix)
    and reconstruction of all k components."""
    x, L, k = args
    import oracle as O
    s = O.RefSSA(np.asarray(x, dtype=np.float64), L).decompose_randomized(k, P_OVERSAMPLE)
    s.reconstruct(list(range(s.k)))
    return len(x)


def _mkl_ssa(args, q=Q_POWER, p=P_OVERSAMPLE):
    """One SSA of series i through the reference's MKL call sequence
    (oracle/mkl_baseline.c): decompose with q power iterations and oversampling
    p, the series' Philox test matrix, reconstruct all k components."""
    x, L, k = args
    import oracle as O
    x = np.asarray(x, dtype=np.float64)
    K = len(x) - L + 1
    b = min(k + p, L, K)
    s = O.MklSSA(x, L).decompose(k, p, q, O.philox_gaussian(4, 0, K, b))
    s.reconstruct(list(range(k)))
    return len(x)


def _mkl_faithful(args):
    """The reference's own shape: q = 0 (no power iterations), p = 8."""
    return _mkl_ssa(args, q=0, p=8)


def _mkl_worker(w, workers, xs, Ls, ks, budget_s, fn, threads):
    """_oracle_worker over the MKL path with `threads` MKL threads, after one
    untimed call (MKL's first call sets up its thread pool and tables)."""
    import oracle as O
    O.mkl_lib().mklb_set_threads(threads)
    fn((xs[w % len(xs)], int(Ls[w % len(xs)]), int(ks[w % len(xs)])))
    return _oracle_worker(w, workers, xs, Ls, ks, budget_s, fn)


def cpu_mkl(xs, Ls, ks, budget_s, workers):
    """The reference's MKL call sequence (decompose_randomized + reconstruct,
    MKL/ssa_opt.h:910-1004, :1685-1763, :2325-2400), linked at run time to
    the MKL runtime on the host, timed on the same batch:
      matched:  q = 2, p = 10 (the device's work), one process per core with
                sequential MKL (series i to process i mod cores), and one
                process with MKL threaded over every core, one series at a
                time (how the reference is used);
      faithful: q = 0, p = 8 (the reference's own shape), one process per core.
    Every leg runs in forked children, so this process never loads MKL.
    None when the MKL runtime is absent."""
    import multiprocessing as mp
    import oracle as O
    if not (os.path.exists(O.MKL_SO) and os.path.exists(O.MKL_RT)):
        return None
    ctx = mp.get_context("fork")
    with ctx.Pool(1) as pool:
        version = pool.apply(O.mkl_version)
        d1, n1, t1 = pool.apply(_mkl_worker, (0, 1, xs, Ls, ks, budget_s, _mkl_ssa, 1))
        dt, nt, tt = pool.apply(_mkl_worker, (0, 1, xs, Ls, ks, budget_s, _mkl_ssa, workers))
    out = {}
    for name, fn in (("matched", _mkl_ssa), ("faithful", _mkl_faithful)):
        with ctx.Pool(workers) as pool:
            res = pool.starmap(_mkl_worker, [(w, workers, xs, Ls, ks, budget_s, fn, 1) for w in range(workers)])
        out[name] = (sum(r[0] for r in res), sum(r[1] for r in res), max(r[2] for r in res))
    dm, nm, em = out["matched"]
    df, nf, ef = out["faithful"]
    return {"value": round(dm / em, 2), "unit": "SSA/s", "cores": workers, "kind": "port",
            "sample": f"the reference's MKL call sequence (oracle/mkl_baseline.c, {version}), FP64: {dm} series of "
                      f"the same batch (sum N = {nm}) over {workers} processes with sequential MKL (series i to "
                      f"process i mod {workers}), decompose q=2 p=10 + reconstruct all k, {em:.1f} s",
            "value_1_core": round(d1 / t1, 2),
            "value_threaded": round(dt / tt, 2),
            "threaded_sample": f"one process, MKL on {workers} threads, one series at a time: {dt} series in {tt:.1f} s",
            "faithful": {"value": round(df / ef, 2),
                         "sample": f"q=0 p=8 (the reference's shape), {df} series over {workers} processes, {ef:.1f} s"}}


def _mkl_configs_child(threads):
    """Per-config latencies of the MKL call sequence (run in a forked child):
    cfg1 (N = 1000, L = 250, k = 50: decompose q=2 p=10 + reconstructions [0]
    and [1, 2] + W-correlation), cfg2 (N = 20000, L = 5000, k = 50: decompose
    + the 50 single-component reconstructions + W-correlation) and the
    reference's published shape (N = 500, L = 375, k = 15, p = 10, q = 0:
    decompose, and decompose + reconstruction of all 15).  One context per
    shape; the test matrix is drawn once, outside the timed region.  p50/p99
    in ms after 3 untimed calls."""
    import oracle as O
    O.mkl_lib().mklb_set_threads(threads)
    out = {}

    def timed(fn, reps):
        for _ in range(3):
            fn()
        t = []
        for _ in range(reps):
            a = time.perf_counter()
            fn()
            t.append(time.perf_counter() - a)
        return round(1e3 * float(np.percentile(t, 50)), 3), round(1e3 * float(np.percentile(t, 99)), 3)

    def shape(N, L, k, p, q, seed, om_seed):
        x = workloads.single(N, seed=seed)
        K = N - L + 1
        return O.MklSSA(x, L), O.philox_gaussian(om_seed, 0, K, min(k + p, L, K))

    m, om = shape(1000, 250, 50, P_OVERSAMPLE, Q_POWER, 1, 11)

    def cfg1():
        m.decompose(50, P_OVERSAMPLE, Q_POWER, om)
        m.reconstruct([0])
        m.reconstruct([1, 2])
        m.wcorr()
    out["cfg1_p50_ms"], out["cfg1_p99_ms"] = timed(cfg1, 20)
    m2, om2 = shape(20000, 5000, 50, P_OVERSAMPLE, Q_POWER, 2, 12)

    def cfg2():
        m2.decompose(50, P_OVERSAMPLE, Q_POWER, om2)
        for i in range(50):
            m2.reconstruct([i])
        m2.wcorr()
    out["cfg2_p50_ms"], out["cfg2_p99_ms"] = timed(cfg2, 10 if threads > 1 else 3)
    m5, om5 = shape(500, 375, 15, P_OVERSAMPLE, 0, 15, 15)
    out["ref_shape_N500_k15_q0_decompose_p50_ms"], _ = timed(lambda: m5.decompose(15, P_OVERSAMPLE, 0, om5), 200)

    def pipe():
        m5.decompose(15, P_OVERSAMPLE, 0, om5)
        m5.reconstruct(list(range(15)))
    out["ref_shape_N500_k15_q0_pipeline_p50_ms"], _ = timed(pipe, 200)
    # the reference's published table: decompose (k = 50, p = 8, q = 0) +
    # reconstruction of all 50, L = N // 4
    for N in README_SHAPES_MS:
        mN, omN = shape(N, N // 4, 50, 8, 0, 20 + N, 21)

        def table_call(m_=mN, om_=omN):
            m_.decompose(50, 8, 0, om_)
            m_.reconstruct(list(range(50)))
        out[f"published_table_N{N}_p50_ms"], _ = timed(table_call, 5)
    return out


def cpu_mkl_configs(workers):
    """_mkl_configs_child with MKL threaded over every core (how the reference
    is used) and on one core, each in a forked child; None without MKL."""
    import multiprocessing as mp
    import oracle as O
    if not (os.path.exists(O.MKL_SO) and os.path.exists(O.MKL_RT)):
        return None
    ctx = mp.get_context("fork")
    with ctx.Pool(1) as pool:
        allc = pool.apply(_mkl_configs_child, (workers,))
    with ctx.Pool(1) as pool:
        one = pool.apply(_mkl_configs_child, (1,))
    return {"threads": workers, "all_cores": allc, "one_core": one,
            "sample": "the reference's MKL call sequence (oracle/mkl_baseline.c), FP64, one prepared context "
                      "per shape, test matrix drawn outside the timed region"}


def _oracle_worker(w, workers, xs, Ls, ks, budget_s, fn=None):
    fn = fn or _oracle_ssa
    done = nsum = 0
    t0 = time.perf_counter()
    for i in range(w, len(xs), workers):
        nsum += fn((xs[i], int(Ls[i]), int(ks[i])))
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    return done, nsum, time.perf_counter() - t0


def cpu_baseline(xs, Ls, ks, budget_s, workers):
    """The FP64 oracle (a C restatement of the reference's call sequence) on
    the same batch: one thread, then one process per core (forked before any
    GPU call), each over its share of the series until budget_s."""
    import multiprocessing as mp
    d1, n1, t1 = _oracle_worker(0, 1, xs, Ls, ks, budget_s)
    one = done_all = nsum_all = 0
    one = d1 / t1
    ctx = mp.get_context("fork")
    with ctx.Pool(workers) as pool:
        res = pool.starmap(_oracle_worker, [(w, workers, xs, Ls, ks, budget_s) for w in range(workers)])
    done_all = sum(r[0] for r in res)
    nsum_all = sum(r[1] for r in res)
    el = max(r[2] for r in res)
    return {"value": round(done_all / el, 2), "unit": "SSA/s", "cores": workers, "kind": "port",
            "sample": f"{done_all} series of the same batch (sum N = {nsum_all}) over {workers} processes "
                      f"(series i to process i mod {workers}), decompose q=2 p=10 + reconstruct all k, FP64, "
                      f"{el:.1f} s; one thread: {one:.2f} SSA/s over the first {d1} series ({n1} samples)",
            "value_1_core": round(one, 2)}


def cpu_reference(xs, Ls, ks, budget_s, workers):
    """The reference's own CPU code (MKL/ssa_opt_ref.h, built from its source
    into oracle/_ref) timed the same way: q = 0 decompose + reconstruct all k,
    FP64.  Its randomized SVD returns a wrong basis (a sign error in its
    Householder QR, MKL/ssa_opt_ref.h:441-519); it is timed for the work it
    does.  None when the build is absent."""
    import multiprocessing as mp
    import oracle as O
    if not O.ref_available():
        return None
    d1, n1, t1 = _oracle_worker(0, 1, xs, Ls, ks, budget_s, _ref_ssa)
    ctx = mp.get_context("fork")
    with ctx.Pool(workers) as pool:
        res = pool.starmap(_oracle_worker, [(w, workers, xs, Ls, ks, budget_s, _ref_ssa) for w in range(workers)])
    done_all = sum(r[0] for r in res)
    el = max(r[2] for r in res)
    return {"value": round(done_all / el, 2), "unit": "SSA/s", "cores": workers, "kind": "reference",
            "sample": f"{done_all} series of the same batch over {workers} processes, the reference's "
                      f"ssa_opt_ref_decompose_randomized (q=0, p=10) + reconstruct all k, FP64, {el:.1f} s; "
                      f"one thread: {d1 / t1:.2f} SSA/s over the first {d1} series ({n1} samples)",
            "value_1_core": round(d1 / t1, 2)}


SERVE_BUDGET_GIB = 32
SERVE_INFLIGHT = 4      # waves in flight: one stream each, as many as the device's hardware queues
SERVE_BACKLOG = 8192    # requests of the saturated backlog the capacity is measured on
SERVE_FIXED_RATE = 20000.0   # req/s: a second open-loop stream at this rate, for comparison across builds


def _oracle_request(args):
    """One serving request by the FP64 oracle: decompose (q = 2, p = 10, the
    request's Philox key), reconstruct the group 0..g-1 and R-forecast h = 50."""
    x, L, k, g, key = args
    import oracle as O
    x = np.asarray(x, dtype=np.float64)
    K = len(x) - L + 1
    b = min(k + P_OVERSAMPLE, L, K)
    o = O.OracleSSA(x, L).decompose(k, P_OVERSAMPLE, Q_POWER, O.philox_gaussian(key, 0, K, b))
    o.reconstruct(list(range(g)))
    if o.lrf(list(range(g)))[0] == 0:
        o.forecast(list(range(g)), 50)
    return len(x)


def _serve_worker(w, workers, reqs, budget_s):
    done = 0
    t0 = time.perf_counter()
    for i in range(w, len(reqs), workers):
        _oracle_request(reqs[i])
        done += 1
        if time.perf_counter() - t0 > budget_s:
            break
    return done, time.perf_counter() - t0


def cpu_serving_baseline(budget_s, workers):
    """The serving workload's CPU line: the FP64 oracle over the same request
    distribution (N log-uniform in [500, 50000]), one process per core, each
    over its share of the requests until budget_s.  The CPU cannot keep up
    with the device's arrival rate, so this is its saturation throughput."""
    import multiprocessing as mp
    import serve
    reqs = [(r.x, r.L, r.k, r.g, serve.request_key(r.rid, 6)) for r in serve.synthetic_requests(2048, rate=np.inf, seed=6)]
    with mp.get_context("fork").Pool(workers) as pool:
        res = pool.starmap(_serve_worker, [(w, workers, reqs, budget_s) for w in range(workers)])
    done = sum(r[0] for r in res)
    el = max(r[1] for r in res)
    return {"capacity_req_s": round(done / el, 2), "cores": workers, "kind": "port",
            "sample": f"{done} requests of the serving distribution over {workers} processes, {el:.1f} s"}


def serving(seconds):
    """cfg5: requests with N log-uniform in [500, 50000], L = N // 4,
    k ~ U{10..50}, reconstruct + forecast h = 50 of components
    0..floor(2k/5)-1, mixed form.  Capacity from a saturated backlog of
    SERVE_BACKLOG requests; then an open-loop Poisson stream at 70 % of it for
    `seconds`, and one at the fixed rate SERVE_FIXED_RATE (the rate earlier
    measurements of this loop were quoted at); then the same kind of request
    served alone.  SERVE_INFLIGHT
    waves are in flight, each a one-branch batch on its own stream; they are
    admitted against a device-memory budget of SERVE_BUDGET_GIB (an equal share
    per in-flight slot)."""
    import serve
    srv = serve.Server(max_wave=512, inflight=SERVE_INFLIGHT, hbm_budget=SERVE_BUDGET_GIB * 2 ** 30)
    # grow the in-flight slots to the largest wave before anything is timed
    srv.warm(sorted(serve.synthetic_requests(4096, rate=np.inf, seed=63), key=lambda r: -len(r.x))[:512])
    done = srv.run(serve.synthetic_requests(SERVE_BACKLOG, rate=np.inf, seed=6))
    capacity = len(done) / max(r.done for r in done)
    rate = 0.7 * capacity
    count = max(64, int(rate * seconds))
    stream = srv.run(serve.synthetic_requests(count, rate=rate, seed=61))
    fixed = srv.run(serve.synthetic_requests(max(64, int(SERVE_FIXED_RATE * seconds)), rate=SERVE_FIXED_RATE,
                                             seed=65)) if SERVE_FIXED_RATE < capacity else None
    alone = []
    one = serve.Server(max_wave=1, inflight=1)
    one.warm(sorted(serve.synthetic_requests(512, rate=np.inf, seed=64), key=lambda r: -len(r.x))[:1])
    for r in serve.synthetic_requests(48, rate=np.inf, seed=62):
        alone.extend(one.run([r]))
    mix, solo = serve.latency_stats(stream), serve.latency_stats(alone)
    srv_wave_bytes = list(srv.wave_bytes)
    del srv, one
    # the plain FP64 form on the same requests: its capacity, then an open-loop
    # stream at the mixed form's rate when it can sustain that, else at 70 %
    # of its own capacity
    psrv = serve.Server(max_wave=512, inflight=SERVE_INFLIGHT, precision=0, hbm_budget=SERVE_BUDGET_GIB * 2 ** 30)
    psrv.warm(sorted(serve.synthetic_requests(4096, rate=np.inf, seed=63), key=lambda r: -len(r.x))[:512])
    pdone = psrv.run(serve.synthetic_requests(SERVE_BACKLOG, rate=np.inf, seed=6))
    pcap = len(pdone) / max(r.done for r in pdone)
    prate = rate if rate <= 0.7 * pcap else 0.7 * pcap
    pstream = psrv.run(serve.synthetic_requests(max(64, int(prate * seconds)), rate=prate, seed=61))
    plain = {"capacity_req_s": round(pcap, 1), "rate_req_s": round(prate, 1),
             "rate_is": "the mixed form's rate" if prate == rate else "70 % of its own capacity",
             "in_mix": serve.latency_stats(pstream)}
    return {"workload": "N log-uniform in [500, 50000], L = N // 4, k ~ U{10..50}, group 0..floor(2k/5)-1, "
                        "h = 50, Poisson arrivals, mixed form (plain FP64 beside it)",
            "hbm_budget_GiB": SERVE_BUDGET_GIB, "max_wave_GB": round(max(srv_wave_bytes) / 1e9, 2),
            "inflight_waves": SERVE_INFLIGHT, "branches_per_wave": 1, "capacity_backlog": SERVE_BACKLOG,
            "capacity_req_s": round(capacity, 1), "rate_req_s": round(rate, 1),
            "in_mix": mix, "alone": solo,
            "at_fixed_rate": None if fixed is None else dict(rate_req_s=SERVE_FIXED_RATE, **serve.latency_stats(fixed)),
            "p50_in_mix_over_alone": round(mix["p50_ms"] / solo["p50_ms"], 3),
            "plain_fp64": plain}


def dense_errors():
    """Error of both forms against a dense FP64 LAPACK SVD of the explicit
    Hankel matrix (numpy) at N = 4000, L = 1000, k = 50, p = 10, q = 2: every
    singular value; the reconstructions of the groups [0], [1, 2], [3, 4],
    0..9 and 0..49 (each dense one by diagonal averaging of sum sigma_i u_i v_i^T,
    as the convolution of u_i and v_i over the anti-diagonal counts); the
    50-step R-forecast of the group 0..4 (LRF from the dense U,
    MKL/ssa_opt.h:2766-2841).  Relative to RMS(x) for reconstructions and
    forecasts.  The q = 2 randomized SVD itself differs from the dense one on
    the noise-level components, equally in both forms."""
    N, L, k = 4000, 1000, 50
    K = N - L + 1
    xd = workloads.single(N, seed=3)
    H = np.lib.stride_tricks.sliding_window_view(xd, K)[:L]
    Ud, sd, Vtd = np.linalg.svd(H, full_matrices=False)
    cnt = np.convolve(np.ones(L), np.ones(K))
    rms = float(np.sqrt(np.mean(xd ** 2)))

    def rec(g):
        return sum(sd[i] * np.convolve(Ud[:, i], Vtd[i]) for i in g) / cnt

    def lrf_forecast(U, base, g, h):
        pi = U[L - 1, g]
        R = U[:L - 1, g] @ pi / (1.0 - float(pi @ pi))
        buf = list(base[N - (L - 1):])
        for _ in range(h):
            buf.append(float(np.dot(R, buf[-(L - 1):])))
        return np.array(buf[L - 1:])

    groups = {"0": [0], "1-2": [1, 2], "3-4": [3, 4], "0-9": list(range(10)), "0-49": list(range(k))}
    g5 = [0, 1, 2, 3, 4]
    ref_f = lrf_forecast(Ud, rec(g5), g5, 50)
    res = {}
    for prec, name in ((0, "plain"), (1, "mixed")):
        s = ssa_amd.SSA(xd, L, q=Q_POWER, precision=prec, seed=13).decompose(k, oversampling=P_OVERSAMPLE)
        sv = s.singular_values
        rel = np.abs(sv - sd[:k]) / sd[:k]
        res[f"dense_sigma_err_{name}"] = float(rel.max())
        res[f"dense_sigma_err_{name}_top10"] = float(rel[:10].max())
        res[f"dense_recon_err_{name}"] = {gn: float(np.abs(s.reconstruct(g) - rec(g)).max() / rms)
                                          for gn, g in groups.items()}
        res[f"dense_forecast_err_{name}"] = float(np.abs(s.forecast(g5, 50) - ref_f).max() / rms)
    return res


def single_configs(reps=3):
    """cfg1: one series N = 1000, L = 250, k = 50: decompose + the
    reconstructions [0] and [1, 2] + W-correlation, plain and mixed.
    cfg2: one series N = 20000, L = 5000, k = 50, p = 10, q = 2:
    decompose + the 50 single-component reconstructions + W-correlation
    (median of `reps`), plain and mixed, and the singular-value error of both
    forms against a dense SVD of the explicit Hankel matrix at N = 4000,
    L = 1000.  cfg4: MSSA M = 16, N = 10000, L = 2500, k = 64, p = 10:
    decompose + per-channel reconstruction of components 0..8.  Streaming
    tick latency (p50, p99) at N = 1000 and N = 20000."""
    out = {}
    # cfg1: N = 1000, L = 250, k = 50: decompose + reconstructions [0] and
    # [1, 2] + the 50 x 50 W-correlation
    x1 = workloads.single(1000, seed=1)
    for prec, name in ((0, "plain"), (1, "mixed")):
        s = ssa_amd.SSA(x1, 250, q=Q_POWER, precision=prec, seed=11)
        t = []
        for r in range(reps + 3):
            a = time.perf_counter()
            s.decompose(50, oversampling=P_OVERSAMPLE)
            s.reconstruct([0])
            s.reconstruct([1, 2])
            s.wcorr_matrix()
            if r >= 2:
                t.append(time.perf_counter() - a)
        out[f"cfg1_{name}_ms"] = round(1e3 * float(np.median(t)), 3)
    x = workloads.single(20000, seed=2)
    groups = [[i] for i in range(50)]
    for prec, name in ((0, "plain"), (1, "mixed")):
        s = ssa_amd.SSA(x, 5000, q=Q_POWER, precision=prec, seed=12)
        t = []
        for r in range(reps + 1):
            a = time.perf_counter()
            s.decompose(50, oversampling=P_OVERSAMPLE)
            s.reconstruct_many(groups)
            s.wcorr_matrix()
            if r:
                t.append(time.perf_counter() - a)
        out[f"cfg2_{name}_ms"] = round(1e3 * float(np.median(t)), 2)
        # the same SSA with the component spectra cached first (ssa_opt_cache_ffts):
        # the 50 reconstructions and the W-correlation read the cached spectra
        t = []
        for r in range(reps + 1):
            a = time.perf_counter()
            s.decompose(50, oversampling=P_OVERSAMPLE)
            s.cache_ffts()
            s.reconstruct_many(groups)
            s.wcorr_matrix_fast()
            if r:
                t.append(time.perf_counter() - a)
        out[f"cfg2_{name}_cached_ms"] = round(1e3 * float(np.median(t)), 2)
    out.update(dense_errors())
    X = workloads.mssa_channels(M=16, N=10000, seed=5)
    for prec, name in ((0, "plain"), (1, "mixed")):
        m = ssa_amd.MSSA(X, 2500, q=Q_POWER, precision=prec)
        t = []
        for r in range(reps + 1):
            a = time.perf_counter()
            m.decompose(64, oversampling=P_OVERSAMPLE)
            for ch in range(16):
                m.reconstruct(ch, list(range(9)))
            if r:
                t.append(time.perf_counter() - a)
        out[f"cfg4_mssa_{name}_ms"] = round(1e3 * float(np.median(t)), 2)
    # the reference's one published latency point (MKL/ssa_opt_r2c.h:9-12: N = 500,
    # k = 15, randomized decompose ~210 us, decompose + reconstruct ~250 us on an
    # i9-14900KF; its harness MKL/test/ssa_block_test.c:426-436 uses L = 375, p = 10):
    # the same calls on one prepared plain FP64 context, p50 over 200 calls, at
    # q = 0 (the reference's algorithm) and q = 2 (this library's default); and
    # the mixed form (FP32 bulk) at q = 0, with its largest relative sigma
    # difference from the plain form's
    x5 = workloads.single(500, seed=15)
    sv_plain = None
    for q, prec in ((0, 0), (2, 0), (0, 1)):
        tag = f"q{q}" + ("_mixed" if prec else "")
        s = ssa_amd.SSA(x5, 375, q=q, precision=prec, seed=15)
        s.decompose(15, oversampling=P_OVERSAMPLE)
        dec, pipe = [], []
        for r in range(200):
            a = time.perf_counter()
            s.decompose(15, oversampling=P_OVERSAMPLE)
            b_ = time.perf_counter()
            s.reconstruct(list(range(15)))
            c = time.perf_counter()
            dec.append(b_ - a)
            pipe.append(c - a)
        out[f"ref_shape_N500_k15_{tag}_decompose_p50_ms"] = round(1e3 * float(np.percentile(dec, 50)), 4)
        out[f"ref_shape_N500_k15_{tag}_pipeline_p50_ms"] = round(1e3 * float(np.percentile(pipe, 50)), 4)
        if q == 0 and not prec:
            sv_plain = s.singular_values.copy()
        elif prec:
            out["ref_shape_N500_k15_q0_mixed_sigma_rel_err"] = float(np.max(np.abs(s.singular_values - sv_plain) / sv_plain[0]))
    # the reference's published single-series table (its README.md:49-63, method
    # MKL/py/benchmark_vs_rssa.py:78-112): decompose_randomized(k = 50, p = 8) +
    # reconstruct(all 50), L = N // 4, median of 5 calls on an i9-14900KF with
    # MKL; here the same calls on one prepared plain FP64 context per N, q = 0
    # (the reference's algorithm), median of 5 after 2 untimed calls
    table = {}
    for N in README_SHAPES_MS:
        xs_ = workloads.single(N, seed=20 + N)
        s = ssa_amd.SSA(xs_, N // 4, q=0, precision=0, seed=21)
        g50 = list(range(50))
        t = []
        for r in range(7):
            a = time.perf_counter()
            s.decompose(50, oversampling=8)
            s.reconstruct(g50)
            if r >= 2:
                t.append(time.perf_counter() - a)
        table[str(N)] = {"ms": round(1e3 * float(np.median(t)), 3), "published_ms": README_SHAPES_MS[N]}
        del s
    out["published_table_q0_p8_k50"] = table
    # streaming ticks, the reference's low-latency pattern (MKL/ssa_opt_r2c.h:82-91):
    # one prepared context; per tick update_signal on the window moved by one
    # sample, decompose_randomized (k = 50, p = 10, q = 2) and the trend
    # reconstruction, timed end to end on the host (plain FP64)
    # "lean": the same ticks with the host U/V mirrors off (ssa_amd_set_mirror),
    # so a decomposition does not copy U and V to the host
    for lean in (False, True):
        for N, L in ((1000, 250), (20000, 5000)):
            xs_long = workloads.single(N + 101, seed=7)
            s = ssa_amd.SSA(xs_long[:N], L, q=Q_POWER, precision=0, seed=14)
            if lean:
                s.set_mirror(False)
            s.decompose(50, oversampling=P_OVERSAMPLE)
            lat = []
            for tick in range(1, 101):
                a = time.perf_counter()
                s.update_signal(xs_long[tick:tick + N])
                s.decompose(50, oversampling=P_OVERSAMPLE)
                s.reconstruct([0])
                lat.append(time.perf_counter() - a)
            tag = "stream_lean" if lean else "stream"
            out[f"{tag}_N{N}_p50_ms"] = round(1e3 * float(np.percentile(lat, 50)), 3)
            out[f"{tag}_N{N}_p99_ms"] = round(1e3 * float(np.percentile(lat, 99)), 3)
    return out


def wcorr_workload(reps=10):
    """A workload for counter runs of the W-correlation Gram (the weighted
    SYRK of MKL/ssa_opt.h:2473): cfg2 (N = 20000, L = 5000, k = 50, p = 10,
    q = 2), one decomposition per form followed by `reps` W-correlations.  The
    decomposition's own Gram launches come first (8 plain, 6 mixed); each
    W-correlation then launches one Gram over the N x k weighted panel."""
    global ssa_amd
    import ssa_amd as _ssa_amd
    ssa_amd = _ssa_amd
    x = workloads.single(20000, seed=2)
    for prec, name in ((0, "plain"), (1, "mixed")):
        s = ssa_amd.SSA(x, 5000, q=Q_POWER, precision=prec, seed=12)
        s.decompose(50, oversampling=P_OVERSAMPLE)
        t = []
        for _ in range(reps):
            a = time.perf_counter()
            s.wcorr_matrix()
            t.append(time.perf_counter() - a)
        print(f"[bench] wcorr {name}: median {1e3 * float(np.median(t)):.3f} ms over {reps}", flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--n-lo", type=int, default=500, help="shortest series of the batch (profiling other mixes)")
    ap.add_argument("--n-hi", type=int, default=20000, help="longest series of the batch")
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--serve-seconds", type=float, default=5.0)
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-configs", action="store_true")
    ap.add_argument("--no-tune", action="store_true", help="plan with the cost model instead of measured costs")
    ap.add_argument("--wcorr-only", action="store_true",
                    help="profiling workload: one cfg2 decomposition per form, then 10 W-correlations each")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this SSA path runs on one GPU; --gpus {args.gpus} is not supported "
              f"(series of a batch would shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    if args.wcorr_only:
        wcorr_workload()
        return

    xs, Ns, Ls, ks = workloads.ragged_batch(count=args.count, n_lo=args.n_lo, n_hi=args.n_hi, seed=4,
                                            dtype=np.float32)
    print(f"[bench] batch: {len(xs)} series, sum N = {int(Ns.sum())}", flush=True)

    # CPU baseline first: its worker processes are forked before any GPU call
    workers = max(1, min(16, int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))))
    cpu = None if args.no_cpu else cpu_baseline(xs, Ls, ks, args.cpu_seconds, workers)
    cpu_ref = None if args.no_cpu else cpu_reference(xs, Ls, ks, args.cpu_seconds, workers)
    cpu_mkl_out = None if args.no_cpu else cpu_mkl(xs, Ls, ks, args.cpu_seconds, workers)
    cpu_cfg = None if (args.no_cpu or args.no_configs) else cpu_mkl_configs(workers)
    if cpu_cfg:
        print(f"[bench] cpu MKL per config ({workers} threads): {cpu_cfg['all_cores']}; one core: "
              f"{cpu_cfg['one_core']}", flush=True)
    if cpu_mkl_out:
        print(f"[bench] cpu MKL (reference call sequence): {cpu_mkl_out['value']} SSA/s on {workers} cores "
              f"(q=2), {cpu_mkl_out['value_1_core']} on one, {cpu_mkl_out['value_threaded']} MKL-threaded; "
              f"q=0 p=8 {cpu_mkl_out['faithful']['value']} SSA/s", flush=True)
    if cpu:
        print(f"[bench] cpu baseline: {cpu['value']} SSA/s on {workers} cores, {cpu['value_1_core']} on one", flush=True)
    if cpu_ref:
        print(f"[bench] reference cpu (ssa_opt_ref.h, q=0): {cpu_ref['value']} SSA/s on {workers} cores, "
              f"{cpu_ref['value_1_core']} on one", flush=True)
    cpu_serve = None if (args.no_cpu or args.no_serve) else cpu_serving_baseline(args.cpu_seconds, workers)
    if cpu_serve:
        print(f"[bench] cpu serving: {cpu_serve['capacity_req_s']} req/s on {workers} cores", flush=True)

    global ssa_amd
    import ssa_amd as _ssa_amd
    ssa_amd = _ssa_amd
    tuned = None
    if not args.no_tune:
        # measured transform costs for the planner (both forms)
        tuned = {"mixed_ns": ssa_amd.autotune(1), "plain_ns": ssa_amd.autotune(0)}
        print(f"[bench] autotune (ns per transform, 2^6..2^13): {[round(v, 2) for v in tuned['mixed_ns'].values()]}",
              flush=True)

    mixed = ssa_amd.Batch(xs, Ls, ks, p=P_OVERSAMPLE, q=Q_POWER, precision=1, seed=4)
    plans = [mixed.plan(i) for i in range(len(xs))]
    total, lat = timed_steps(mixed, args.steps, args.warmup)
    ssa_s = len(xs) * args.steps / total
    print(f"[bench] mixed: {ssa_s:.0f} SSA/s, {1e3 * total / args.steps:.2f} ms/step", flush=True)
    checked = step_check(mixed)
    roof = roofline_of(mixed, plans, Ns, Ls, 4)
    alg = sum(algorithmic_bytes(int(Ns[i]), int(Ls[i]), plans[i]["k"], plans[i]["b"], 4) for i in range(len(xs)))
    del mixed

    plain = checked_plain = None
    if not args.no_plain:
        pb = ssa_amd.Batch(xs, Ls, ks, p=P_OVERSAMPLE, q=Q_POWER, precision=0, seed=4)
        ptot, plat = timed_steps(pb, args.plain_steps, 1)
        plain = len(xs) * args.plain_steps / ptot
        checked_plain = step_check(pb)
        print(f"[bench] plain FP64: {plain:.0f} SSA/s, {1e3 * ptot / args.plain_steps:.2f} ms/step", flush=True)
        del pb

    serving_out = None
    if not args.no_serve:
        serving_out = serving(args.serve_seconds)
        print(f"[bench] serving: capacity {serving_out['capacity_req_s']} req/s, at {serving_out['rate_req_s']} "
              f"req/s p50 {serving_out['in_mix']['p50_ms']} ms p99 {serving_out['in_mix']['p99_ms']} ms; alone "
              f"p50 {serving_out['alone']['p50_ms']} ms", flush=True)
        if serving_out["at_fixed_rate"]:
            fx = serving_out["at_fixed_rate"]
            print(f"[bench] serving at {fx['rate_req_s']:.0f} req/s: p50 {fx['p50_ms']} ms p99 {fx['p99_ms']} ms",
                  flush=True)
        pl = serving_out["plain_fp64"]
        print(f"[bench] serving plain FP64: capacity {pl['capacity_req_s']} req/s, at {pl['rate_req_s']} req/s "
              f"({pl['rate_is']}) p50 {pl['in_mix']['p50_ms']} ms p99 {pl['in_mix']['p99_ms']} ms", flush=True)


    configs_out = None
    if not args.no_configs:
        configs_out = single_configs()
        print(f"[bench] single series: {configs_out}", flush=True)

    ms_step = 1e3 * total / args.steps
    out = {
        "metric": "SSA/sec batched (N=500-20k, k=50) + p50/p99 latency; mixed vs FP64 plain",
        "value": round(ssa_s, 1),
        "unit": "SSA/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {"workload": f"ragged batch: {len(xs)} series, N log-uniform in [{args.n_lo}, {args.n_hi}], L = N // 4, "
                               f"k ~ U{{10..50}}, p = 10, q = 2, FP32 inputs; one step = spectra + "
                               f"decomposition + reconstruction of all k components of every series",
                   "series": len(xs), "sum_N": int(Ns.sum()), "form": "mixed (FP32 bulk, FP64 small solves)"},
        "p50_ms_per_step": round(1e3 * float(np.percentile(lat, 50)), 3),
        "p99_ms_per_step": round(1e3 * float(np.percentile(lat, 99)), 3),
        "plain_fp64_value": None if plain is None else round(plain, 1),
        "mixed_over_plain": None if plain is None else round(ssa_s / plain, 3),
        "step_algorithmic_GB": round(alg / 1e9, 3),
        "step_achieved_GBs": round(alg / (ms_step * 1e-3) / 1e9, 1),
    }
    out.update(roof)
    out["output_check"] = {"mixed": checked, "plain": checked_plain}
    out["autotune"] = tuned
    out["serving"] = serving_out
    if serving_out is not None:
        serving_out["cpu_fp64"] = cpu_serve
    out["single_series"] = configs_out
    # the MKL-linked reference call sequence where the MKL runtime is on the
    # host, the FP64 oracle port otherwise; both are reported
    out["cpu_baseline"] = cpu_mkl_out or cpu
    out["cpu_oracle"] = cpu
    out["cpu_configs"] = cpu_cfg
    out["cpu_reference"] = cpu_ref
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
