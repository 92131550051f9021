"""C-ABI conformance of libssa.so (CPU only: no compute call needs a GPU).

* every symbol include/ssa_opt.h declares is exported;
* the reference wrapper's manifest (MKL/py/ssa_wrapper.py:125-291): the 24
  symbols it binds unconditionally, plus the getters, are present;
* sizeof(SSA_Opt), sizeof(MSSA_Opt) fit the wrapper's opaque 1024-byte
  blocks (MKL/py/ssa_wrapper.py:92-98) and the public fields C callers read
  exist (MKL/test/ssa_opt_test.c:176-182);
* the correlation planner covers every (N, L) with n >= CL + CK - 1.
"""
import ctypes
import os
import re
import subprocess
import tempfile

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HDR = os.path.join(ROOT, "include", "ssa_opt.h")
LIB = os.path.join(ROOT, "singular-spectrum-analysis_amd", "libssa.so")

# MKL/py/ssa_wrapper.py:125-204, :272-291 -- a missing one breaks the wrapper's import
WRAPPER_UNCONDITIONAL = [
    "ssa_opt_init", "ssa_opt_decompose", "ssa_opt_decompose_block", "ssa_opt_decompose_randomized",
    "ssa_opt_prepare", "ssa_opt_update_signal", "ssa_opt_free_prepared", "ssa_opt_reconstruct",
    "ssa_opt_forecast", "ssa_opt_forecast_full", "ssa_opt_wcorr_matrix", "ssa_opt_wcorr_pair",
    "ssa_opt_variance_explained", "ssa_opt_find_periodic_pairs", "ssa_opt_get_trend", "ssa_opt_get_noise",
    "ssa_opt_free", "mssa_opt_init", "mssa_opt_decompose", "mssa_opt_reconstruct", "mssa_opt_reconstruct_all",
    "mssa_opt_series_contributions", "mssa_opt_variance_explained", "mssa_opt_free",
]
# MKL/py/ssa_wrapper.py:189-201 (needed by SSA.singular_values / eigenvalues / total_variance)
WRAPPER_GETTERS = ["ssa_opt_get_singular_values", "ssa_opt_get_eigenvalues", "ssa_opt_get_total_variance"]


def declared_functions():
    src = open(HDR).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = re.findall(r"\b([a-z_][a-z0-9_]*)\s*\(", src)
    skip = {"if", "sizeof", "defined"}
    return sorted({n for n in names if (n.startswith("ssa_") or n.startswith("mssa_")) and n not in skip})


@pytest.fixture(scope="module")
def lib():
    if not os.path.exists(LIB):
        pytest.fail("libssa.so is not built (run __graft_entry__.build())")
    return ctypes.CDLL(LIB)


def test_exports_every_declared_symbol(lib):
    missing = [n for n in declared_functions() if not hasattr(lib, n)]
    assert not missing, missing
    assert len(declared_functions()) >= 45


# MKL/py/ssa_wrapper.py:156-168: bound only when present; SSA.vforecast needs them
WRAPPER_VFORECAST = ["ssa_opt_vforecast", "ssa_opt_vforecast_full", "ssa_opt_vforecast_fast"]
# MKL/py/ssa_wrapper.py:209-221, :259-266, :236-246 (optional families) and the
# MKL hooks the reference's benches call (MKL/py/benchmark_vs_rssa.py:33-43)
WRAPPER_OPTIONAL = ["ssa_opt_cadzow", "ssa_opt_cadzow_weighted", "ssa_opt_cadzow_inplace", "ssa_opt_gapfill",
                    "ssa_opt_gapfill_simple", "ssa_opt_parestimate", "ssa_opt_parestimate_free", "ssa_mkl_init"]


def test_mkl_hooks_are_harmless(lib):
    assert lib.ssa_mkl_init(1) == 0
    assert lib.ssa_mkl_get_threads() == 1
    p = ctypes.c_int(-1)
    lib.ssa_mkl_get_cpu_info(ctypes.byref(p), None, None, None)
    assert p.value >= 1


def test_wrapper_manifest(lib):
    for n in WRAPPER_UNCONDITIONAL + WRAPPER_GETTERS + WRAPPER_VFORECAST + WRAPPER_OPTIONAL:
        assert hasattr(lib, n), n
    assert len(WRAPPER_UNCONDITIONAL) == 24


def test_python_mirror_binds_every_symbol():
    import ssa_amd
    bound = {n for n, _, _ in ssa_amd.SIGNATURES}
    decl = set(declared_functions())
    assert decl <= bound, sorted(decl - bound)
    ssa_amd.load()   # binds argtypes/restype of every symbol; raises if one is missing


def test_struct_sizes_and_fields():
    prog = r"""
#include <stdio.h>
#include <stddef.h>
#include "ssa_opt.h"
int main(void) {
    SSA_Opt s; MSSA_Opt m; (void)s; (void)m;
    printf("%zu %zu %zu\n", sizeof(SSA_Opt), sizeof(MSSA_Opt), sizeof(SSA_LRF));
    /* fields read by the reference's C tests */
    printf("%d\n", (int)(offsetof(SSA_Opt, N) + offsetof(SSA_Opt, fft_x) + offsetof(SSA_Opt, ws_real)
        + offsetof(SSA_Opt, ws_complex) + offsetof(SSA_Opt, sigma) + offsetof(SSA_Opt, eigenvalues)
        + offsetof(SSA_Opt, U) + offsetof(SSA_Opt, V) + offsetof(SSA_Opt, n_components)
        + offsetof(SSA_Opt, initialized) + offsetof(SSA_Opt, decomposed) + offsetof(SSA_Opt, prepared)
        + offsetof(SSA_Opt, fft_len) + offsetof(SSA_Opt, r2c_len) + offsetof(SSA_Opt, batch_size)
        + offsetof(SSA_Opt, decomp_Omega) + offsetof(SSA_Opt, decomp_Y) + offsetof(SSA_Opt, ws_batch_real)));
    return 0;
}
"""
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(prog)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HDR), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.split()
    s1, s2, s3 = int(out[0]), int(out[1]), int(out[2])
    assert s1 <= 1024 and s2 <= 1024 and s3 <= 256


def test_python_context_mirrors_match_header():
    """The typed ctypes mirrors of SSA_Opt and MSSA_Opt in ssa_amd.py have the
    header's size and every field at the header's offset."""
    import ssa_amd
    lines = []
    for cname, py in (("SSA_Opt", ssa_amd._SSA_Opt), ("MSSA_Opt", ssa_amd._MSSA_Opt)):
        lines.append(f'    printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in py._fields_:
            lines.append(f'    printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    prog = '#include <stdio.h>\n#include <stddef.h>\n#include "ssa_opt.h"\nint main(void) {\n' + \
        "\n".join(lines) + "\n    return 0;\n}\n"
    with tempfile.TemporaryDirectory() as d:
        c = os.path.join(d, "t.c")
        open(c, "w").write(prog)
        exe = os.path.join(d, "t")
        subprocess.run(["gcc", "-std=c11", "-I", os.path.dirname(HDR), c, "-o", exe], check=True)
        out = subprocess.run([exe], capture_output=True, text=True, check=True).stdout.splitlines()
    got = {(a, b): int(v) for a, b, v in (ln.split() for ln in out)}
    for cname, py in (("SSA_Opt", ssa_amd._SSA_Opt), ("MSSA_Opt", ssa_amd._MSSA_Opt)):
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f, _ in py._fields_:
            assert got[(cname, f)] == getattr(py, f).offset, (cname, f)


class Plan(ctypes.Structure):
    _fields_ = [("n", ctypes.c_int), ("lgn", ctypes.c_int), ("CL", ctypes.c_int), ("CK", ctypes.c_int),
                ("AL", ctypes.c_int), ("SK", ctypes.c_int), ("cost", ctypes.c_double)]


@pytest.mark.parametrize("prec", [0, 1])
def test_planner_covers_shapes(lib, prec):
    f = lib.ssa_plan_series
    f.argtypes = [ctypes.c_int] * 8 + [ctypes.POINTER(Plan)]
    f.restype = ctypes.c_int
    shapes = [(4, 2), (4, 3), (37, 2), (37, 36), (100, 25), (1000, 250), (20000, 5000), (50000, 12500),
              (10000, 2500), (8193, 4096), (9000, 7000), (131072, 32768)]
    for N, L in shapes:
        for M in (1, 16):
            p = Plan()
            assert f(prec, N, L, M, 60, 50, 2, -1, ctypes.byref(p)) == 0
            K = N - L + 1
            assert 64 <= p.n <= 8192 and p.n == 1 << p.lgn
            assert p.CL + p.CK - 1 <= p.n
            assert p.AL * p.CL >= L and (p.AL - 1) * p.CL < L
            assert p.SK * p.CK >= K and (p.SK - 1) * p.CK < K
