"""The CPU legs of bench.py (host only): the FP64 oracle baseline and the
reference's own CPU code, on a few short series with a small time budget."""
import numpy as np
import pytest

import workloads


def _bench():
    import bench
    return bench


@pytest.mark.parametrize("leg", ["cpu_baseline", "cpu_reference"])
def test_cpu_legs_report_a_rate(leg):
    B = _bench()
    import oracle as O
    if leg == "cpu_reference" and not O.ref_available():
        pytest.skip("oracle/_ref not built")
    xs, Ns, Ls, ks = workloads.ragged_batch(count=6, n_lo=500, n_hi=900, seed=4, dtype=np.float32)
    r = getattr(B, leg)(xs, Ls, ks, 0.2, 2)
    assert r["unit"] == "SSA/s" and r["cores"] == 2
    assert r["kind"] == ("port" if leg == "cpu_baseline" else "reference")
    assert r["value"] > 0 and r["value_1_core"] > 0
