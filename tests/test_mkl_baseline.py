"""The MKL-linked CPU baseline (oracle/mkl_baseline.c: the reference's call
sequence MKL/ssa_opt.h:910-1004, :1685-1763, :2325-2400 against MKL's C API,
resolved at run time) computes the same decomposition as the FP64 oracle on
the same test matrix.  This pins the DFTI configuration it uses (real
transforms, conjugate-even storage, batch distances, backward scale) and the
LAPACK/BLAS calls: a wrong enumeration value, layout or scale fails here.

The MKL side runs in a fresh interpreter: this test process has PyTorch's CPU
library loaded (the GPU probe of conftest), whose exported BLAS/LAPACK names
MKL's dynamically loaded layers would bind to.  bench.py times the baseline
in processes that never load PyTorch.  Skipped where the MKL runtime is
absent."""
import os
import subprocess
import sys

import numpy as np
import pytest

import oracle as O
import workloads

pytestmark = pytest.mark.skipif(not (os.path.exists(O.MKL_SO) and os.path.exists(O.MKL_RT)),
                                reason="MKL runtime (libmkl_rt) or the baseline build not on this host")

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

# (N, L, k, p, q): the reference's shape (q = 0, p = 8), the device's (q = 2,
# p = 10), a block narrower than the transform batch (4 + 2), and a length
# whose transform batch (16 columns at n = 65536) is smaller than k + p
CASES = [(1000, 250, 15, 8, 0), (1000, 250, 50, 10, 2), (301, 100, 4, 2, 1), (20000, 5000, 30, 10, 2)]
GROUPS = lambda k: ([0], [1, 2], list(range(k)))  # noqa: E731

_CHILD = r"""
import sys
import numpy as np
sys.path[:0] = [{oracle!r}, {pkg!r}]
import oracle as O
import workloads
out = {{}}
for N, L, k, p, q in {cases!r}:
    x = workloads.single(N, seed=N)
    om = O.philox_gaussian(9, 0, N - L + 1, k + p)
    m = O.MklSSA(x, L).decompose(k, p, q, om)
    out[f"s_{{N}}_{{q}}"] = m.sigma
    out[f"u_{{N}}_{{q}}"] = m.U
    for j, g in enumerate(([0], [1, 2], list(range(k)))):
        out[f"r_{{N}}_{{q}}_{{j}}"] = m.reconstruct(g)
    out[f"w_{{N}}_{{q}}"] = m.wcorr()
np.savez({path!r}, **out)
"""


def rms(x):
    return float(np.sqrt(np.mean(np.asarray(x) ** 2)))


@pytest.fixture(scope="module")
def mkl_results(tmp_path_factory):
    path = str(tmp_path_factory.mktemp("mkl") / "mkl.npz")
    code = _CHILD.format(oracle=os.path.join(ROOT, "oracle"), pkg=os.path.join(ROOT, "singular-spectrum-analysis_amd"),
                         cases=CASES, path=path)
    env = dict(os.environ, MKL_NUM_THREADS="1")
    subprocess.run([sys.executable, "-c", code], check=True, timeout=300, env=env)
    return dict(np.load(path))


@pytest.mark.parametrize("N,L,k,p,q", CASES)
def test_mkl_baseline_matches_oracle(mkl_results, N, L, k, p, q):
    x = workloads.single(N, seed=N)
    om = O.philox_gaussian(9, 0, N - L + 1, k + p)
    o = O.OracleSSA(x, L).decompose(k, p, q, om)
    sig = mkl_results[f"s_{N}_{q}"]
    assert np.max(np.abs(sig - o.sigma) / o.sigma) < 1e-12
    r = rms(x)
    for j, g in enumerate(GROUPS(k)):
        assert np.abs(mkl_results[f"r_{N}_{q}_{j}"] - o.reconstruct(g)).max() / r < 1e-11, g
    assert np.abs(mkl_results[f"w_{N}_{q}"] - o.wcorr()).max() < 1e-10
    # the sign rule: every column of U sums to a non-negative value
    assert np.all(mkl_results[f"u_{N}_{q}"].sum(axis=0) >= 0)
