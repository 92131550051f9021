"""Autotuner: measured per-transform costs replace the planner's model for
the contexts and batches created afterwards (process-wide, so this file runs
after the other GPU tests).  Plans made with measured costs must give the
same answers to the same tolerances."""
import numpy as np
import pytest

import oracle as O
import workloads

pytestmark = pytest.mark.gpu


def test_autotune_then_parity():
    import ssa_amd as S
    ns = S.autotune(1)
    assert sorted(ns) == list(range(6, 14))
    assert all(v > 0 and np.isfinite(v) for v in ns.values())
    assert ns[13] > ns[9]          # an 8192-point transform costs more than a 512-point one
    ns0 = S.autotune(0)
    assert all(v > 0 for v in ns0.values())
    # a mixed batch planned with the measured costs against the FP64 oracle
    xs, Ns, Ls, ks = workloads.ragged_batch(count=6, n_lo=500, n_hi=9000, seed=41, dtype=np.float32)
    b = S.Batch(xs, Ls, ks, p=10, q=2, precision=1, seed=3)
    b.run(use_graph=False)
    b.sync()
    for i in range(len(xs)):
        x = np.asarray(xs[i], dtype=np.float64)
        L, k = int(Ls[i]), int(ks[i])
        K = len(x) - L + 1
        key = (3 ^ (0x9E3779B97F4A7C15 * (i + 1))) & ((1 << 64) - 1)
        orc = O.OracleSSA(x, L).decompose(k, 10, 2, O.philox_gaussian(key, 0, K, min(k + 10, L, K)))
        rec = orc.reconstruct(list(range(k)))
        assert np.linalg.norm(b.output(i) - rec) / np.linalg.norm(rec) < 1e-5
    # a plain context planned with the measured costs keeps 1e-10 parity
    x = workloads.single(4000, seed=8)
    om = O.philox_gaussian(77, 0, 3001, 40)
    dev = S.SSA(x, 1000, q=2, precision=0)
    dev.prepare(30, 10)
    dev.inject_omega(om)
    dev.decompose(30, oversampling=10)
    orc = O.OracleSSA(x, 1000).decompose(30, 10, 2, om)
    assert np.max(np.abs(dev.singular_values - orc.sigma) / orc.sigma) < 1e-10
