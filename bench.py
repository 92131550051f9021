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

import ssa_amd  # noqa: E402
import workloads  # noqa: E402

HBM_PEAK_GBS = 8000.0
P_OVERSAMPLE = 10
Q_POWER = 2


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
    # dominant kernel instantiation: k_matvec<T, lg, 0> (both directions move
    # the same bytes; the generated-Omega launch is another instantiation)
    mv = {}
    for (ph, lg), v in agg.items():
        if ph in ("matvec_H", "matvec_HT"):
            mv.setdefault(lg, []).extend(v)
    lg, v = max(mv.items(), key=lambda kv: sum(kv[1]))
    sel = [i for i, p in enumerate(plans) if p["lgn"] == lg]
    bytes_launch = float(sum(matvec_bytes(int(Ns[i]), int(Ls[i]), plans[i]["b"], s) for i in sel))
    avg_ms = float(np.mean(v))
    achieved = bytes_launch / (avg_ms * 1e-3) / 1e9
    kernel = f"k_matvec<{'float' if s == 4 else 'double'}, {lg}>"
    traffic = measured_traffic(kernel)
    return {
        "roofline": {"bound": "hbm", "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "traffic": None if traffic is None else round(traffic),
                     "kernel": kernel,
                     "bytes_per_launch": bytes_launch, "avg_launch_ms": round(avg_ms, 4),
                     "series_in_launch": len(sel), "launches_timed": len(v)},
        "phase_ms": {k: round(v, 3) for k, v in sorted(phase_tot.items(), key=lambda kv: -kv[1])},
        "profiled_step_ms": round(total, 3),
    }


def cpu_baseline(xs, Ls, ks, budget_s):
    """The FP64 oracle (a C restatement of the reference's call sequence,
    one thread) on the first series of the same batch until budget_s."""
    import oracle as O
    done = 0
    nsum = 0
    t0 = time.perf_counter()
    for i in range(len(xs)):
        x = np.asarray(xs[i], dtype=np.float64)
        L, k = int(Ls[i]), int(ks[i])
        K = len(x) - L + 1
        b = min(k + P_OVERSAMPLE, L, K)
        om = O.philox_gaussian(4, 0, K, b)
        s = O.OracleSSA(x, L).decompose(k, P_OVERSAMPLE, Q_POWER, om)
        s.reconstruct(list(range(k)))
        done += 1
        nsum += len(x)
        if time.perf_counter() - t0 > budget_s:
            break
    el = time.perf_counter() - t0
    return {"value": round(done / el, 2), "unit": "SSA/s", "cores": 1, "kind": "port",
            "sample": f"first {done} series of the same batch (sum N = {nsum}), decompose q=2 p=10 + "
                      f"reconstruct all k, FP64, {el:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--count", type=int, default=4096)
    ap.add_argument("--plain-steps", type=int, default=5)
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-plain", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: this SSA path runs on one GPU; --gpus {args.gpus} is not supported "
              f"(series of a batch would shard across GPUs with no exchange)", file=sys.stderr)
        sys.exit(2)

    xs, Ns, Ls, ks = workloads.ragged_batch(count=args.count, seed=4, dtype=np.float32)
    print(f"[bench] batch: {len(xs)} series, sum N = {int(Ns.sum())}", flush=True)

    mixed = ssa_amd.Batch(xs, Ls, ks, p=P_OVERSAMPLE, q=Q_POWER, precision=1, seed=4)
    plans = [mixed.plan(i) for i in range(len(xs))]
    total, lat = timed_steps(mixed, args.steps, args.warmup)
    ssa_s = len(xs) * args.steps / total
    print(f"[bench] mixed: {ssa_s:.0f} SSA/s, {1e3 * total / args.steps:.2f} ms/step", flush=True)
    roof = roofline_of(mixed, plans, Ns, Ls, 4)
    alg = sum(algorithmic_bytes(int(Ns[i]), int(Ls[i]), plans[i]["k"], plans[i]["b"], 4) for i in range(len(xs)))
    del mixed

    plain = None
    if not args.no_plain:
        pb = ssa_amd.Batch(xs, Ls, ks, p=P_OVERSAMPLE, q=Q_POWER, precision=0, seed=4)
        ptot, plat = timed_steps(pb, args.plain_steps, 1)
        plain = len(xs) * args.plain_steps / ptot
        print(f"[bench] plain FP64: {plain:.0f} SSA/s, {1e3 * ptot / args.plain_steps:.2f} ms/step", flush=True)
        del pb

    cpu = None if args.no_cpu else cpu_baseline(xs, Ls, ks, args.cpu_seconds)
    if cpu:
        print(f"[bench] cpu baseline: {cpu['value']} SSA/s", flush=True)

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
        "config": {"workload": f"ragged batch: {len(xs)} series, N log-uniform in [500, 20000], L = N // 4, "
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
    out["cpu_baseline"] = cpu
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
