"""ssa_amd_reconstruct_noalloc (the allocation-free reconstruction) against
ssa_opt_reconstruct on the same factors, through the C ABI.

The factors come from a plain FP64 context's host mirrors (U, V, sigma),
copied to device buffers; the allocation-free call writes its output into a
device buffer.  Two calls in a row, the second with a smaller series, each
within 1e-9 of RMS(x) of the context's own reconstruction (the plain form's
parity bar); invalid group indices are skipped as in ssa_opt_reconstruct;
sizes beyond the work space and bad arguments are refused without device work.
"""
import ctypes

import numpy as np
import pytest

import workloads

pytestmark = pytest.mark.gpu


def _context(N, L, k):
    import ssa_amd
    x = workloads.single(N, seed=N)
    s = ssa_amd.SSA(x, L, q=2, precision=0, seed=N)
    s.decompose(k, oversampling=10)
    U, V, sig = s.factors()
    return x, s, np.asfortranarray(U).ravel(order="F"), np.asfortranarray(V).ravel(order="F"), sig


def _call(lib, dU, dV, ds, N, L, k, group, out):
    import torch
    g = np.ascontiguousarray(group, dtype=np.int32)
    return lib.ssa_amd_reconstruct_noalloc(dU.data_ptr(), dV.data_ptr(), ds.data_ptr(), N, L, k,
                                           g.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), len(g), out.data_ptr(),
                                           ctypes.c_void_p(torch.cuda.current_stream().cuda_stream))


def test_noalloc_matches_reconstruct():
    import torch
    import ssa_amd
    lib = ssa_amd.lib()
    dev = torch.device("cuda:0")
    for N, L, k, group in ((20000, 5000, 30, [0, 1, 2]), (1000, 250, 12, [0, 3, 99, 1, -1])):
        print(f"noalloc case N={N} L={L} k={k} group={group}", flush=True)
        x, s, U, V, sig = _context(N, L, k)
        ref = s.reconstruct(group)
        dU = torch.from_numpy(U).to(dev)
        dV = torch.from_numpy(V).to(dev)
        ds = torch.from_numpy(sig).to(dev)
        out = torch.full((N,), float("nan"), dtype=torch.float64, device=dev)
        assert _call(lib, dU, dV, ds, N, L, k, group, out) == 0
        torch.cuda.synchronize()
        got = out.cpu().numpy()
        r = float(np.sqrt(np.mean(x ** 2)))
        err = float(np.abs(got - ref).max() / r)
        print(f"  max |noalloc - ssa_opt_reconstruct| / RMS = {err:.2e}", flush=True)
        assert err < 1e-9


def test_noalloc_refusals():
    import torch
    import ssa_amd
    lib = ssa_amd.lib()
    dev = torch.device("cuda:0")
    buf = torch.zeros(16, dtype=torch.float64, device=dev)
    # beyond the work space (N = 400000, L = N / 4) and beyond the group limit: -2
    assert _call(lib, buf, buf, buf, 400000, 100000, 4, [0], buf) == -2
    assert _call(lib, buf, buf, buf, 1000, 250, 4, list(range(300)), buf) == -2
    # invalid arguments: -1
    assert _call(lib, buf, buf, buf, 1000, 250, 4, [], buf) == -1
    assert _call(lib, buf, buf, buf, 1000, 1000, 4, [0], buf) == -1
    assert _call(lib, buf, buf, buf, 1000, 250, 0, [0], buf) == -1
