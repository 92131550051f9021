"""Python front end of the device SSA library (libssa.so).  Its SSA and MSSA
classes offer the interface of the reference's ctypes wrapper
MKL/py/ssa_wrapper.py -- the same class and method names, argument meaning and
error types (ValueError when the library refuses a series, RuntimeError when a
call fails) -- over the same C symbols; they hold the typed C context and read
the library's state from it.  Batch (many series per device step) and the
module functions have no counterpart there.

The library is loaded from this directory; a missing or unloadable library
raises immediately -- there is no CPU fallback.
"""
import ctypes
import os
from typing import List

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libssa.so")

c_double_p = ctypes.POINTER(ctypes.c_double)
c_int_p = ctypes.POINTER(ctypes.c_int)
c_float_p = ctypes.POINTER(ctypes.c_float)
c_ll_p = ctypes.POINTER(ctypes.c_longlong)


BMAX = 80   # widest block k + p the device path takes (SSA_BMAX)

_dpp = ctypes.POINTER(ctypes.c_double)


class _SSA_Opt(ctypes.Structure):
    """SSA_Opt as include/ssa_opt.h lays it out (sizeof below the 1024 bytes the
    reference wrapper reserves, MKL/py/ssa_wrapper.py:92-94)."""
    _fields_ = [("N", ctypes.c_int), ("L", ctypes.c_int), ("K", ctypes.c_int),
                ("fft_len", ctypes.c_int), ("r2c_len", ctypes.c_int), ("batch_size", ctypes.c_int),
                ("fft_x", _dpp), ("ws_real", _dpp), ("ws_complex", _dpp), ("ws_real2", _dpp), ("ws_proj", _dpp),
                ("ws_batch_real", _dpp), ("ws_batch_complex", _dpp),
                ("U", _dpp), ("V", _dpp), ("sigma", _dpp), ("eigenvalues", _dpp), ("n_components", ctypes.c_int),
                ("inv_diag_count", _dpp), ("U_fft", _dpp), ("V_fft", _dpp), ("fft_cached", ctypes.c_bool),
                ("prepared_kp", ctypes.c_int), ("prepared_lwork", ctypes.c_int),
                ("decomp_Omega", _dpp), ("decomp_Y", _dpp),
                ("prepared", ctypes.c_bool), ("initialized", ctypes.c_bool), ("decomposed", ctypes.c_bool),
                ("total_variance", ctypes.c_double), ("dev", ctypes.c_void_p), ("q", ctypes.c_int),
                ("precision", ctypes.c_int), ("seed", ctypes.c_ulonglong), ("calls", ctypes.c_ulonglong)]


class _MSSA_Opt(ctypes.Structure):
    """MSSA_Opt as include/ssa_opt.h lays it out."""
    _fields_ = [("M", ctypes.c_int), ("N", ctypes.c_int), ("L", ctypes.c_int), ("K", ctypes.c_int),
                ("fft_len", ctypes.c_int), ("r2c_len", ctypes.c_int),
                ("U", _dpp), ("V", _dpp), ("sigma", _dpp), ("eigenvalues", _dpp), ("n_components", ctypes.c_int),
                ("inv_diag_count", _dpp), ("initialized", ctypes.c_bool), ("decomposed", ctypes.c_bool),
                ("total_variance", ctypes.c_double), ("dev", ctypes.c_void_p), ("q", ctypes.c_int),
                ("precision", ctypes.c_int), ("seed", ctypes.c_ulonglong), ("calls", ctypes.c_ulonglong)]


class _SSA_LRF(ctypes.Structure):
    _fields_ = [("R", c_double_p), ("L", ctypes.c_int), ("verticality", ctypes.c_double),
                ("valid", ctypes.c_bool)]


class _SSA_CadzowResult(ctypes.Structure):
    # MKL/py/ssa_wrapper.py:108-114
    _fields_ = [("iterations", ctypes.c_int), ("final_diff", ctypes.c_double), ("converged", ctypes.c_double)]


class _SSA_GapFillResult(ctypes.Structure):
    # MKL/py/ssa_wrapper.py:249-255
    _fields_ = [("iterations", ctypes.c_int), ("final_diff", ctypes.c_double), ("converged", ctypes.c_int),
                ("n_gaps", ctypes.c_int)]


class _SSA_ParEstimate(ctypes.Structure):
    # MKL/ssa_opt.h:389-396
    _fields_ = [("periods", c_double_p), ("frequencies", c_double_p), ("moduli", c_double_p),
                ("rates", c_double_p), ("n_components", ctypes.c_int)]


# (name, restype, argtypes) of every symbol include/ssa_opt.h declares
_P = ctypes.POINTER(_SSA_Opt)
_M = ctypes.POINTER(_MSSA_Opt)
_I, _D, _U64, _VP = ctypes.c_int, ctypes.c_double, ctypes.c_ulonglong, ctypes.c_void_p
SIGNATURES = [
    ("ssa_opt_init", _I, [_P, c_double_p, _I, _I]),
    ("ssa_opt_free", None, [_P]),
    ("ssa_opt_prepare", _I, [_P, _I, _I]),
    ("ssa_opt_free_prepared", None, [_P]),
    ("ssa_opt_update_signal", _I, [_P, c_double_p]),
    ("ssa_opt_decompose_randomized", _I, [_P, _I, _I]),
    ("ssa_opt_decompose", _I, [_P, _I, _I]),
    ("ssa_opt_decompose_block", _I, [_P, _I, _I, _I]),
    ("ssa_opt_extend", _I, [_P, _I, _I]),
    ("ssa_opt_hankel_matvec", None, [_P, c_double_p, c_double_p]),
    ("ssa_opt_hankel_matvec_T", None, [_P, c_double_p, c_double_p]),
    ("ssa_opt_reconstruct", _I, [_P, c_int_p, _I, c_double_p]),
    ("ssa_opt_cache_ffts", _I, [_P]),
    ("ssa_opt_free_cached_ffts", None, [_P]),
    ("ssa_opt_wcorr_matrix", _I, [_P, c_double_p]),
    ("ssa_opt_wcorr_matrix_fast", _I, [_P, c_double_p]),
    ("ssa_opt_wcorr_pair", _D, [_P, _I, _I]),
    ("ssa_opt_component_stats", _I, [_P, _VP]),
    ("ssa_opt_component_stats_free", None, [_VP]),
    ("ssa_opt_find_periodic_pairs", _I, [_P, c_int_p, _I, _D, _D]),
    ("ssa_opt_get_trend", _I, [_P, c_double_p]),
    ("ssa_opt_get_noise", _I, [_P, _I, c_double_p]),
    ("ssa_opt_variance_explained", _D, [_P, _I, _I]),
    ("ssa_opt_get_singular_values", _I, [_P, c_double_p, _I]),
    ("ssa_opt_get_eigenvalues", _I, [_P, c_double_p, _I]),
    ("ssa_opt_get_total_variance", _D, [_P]),
    ("ssa_opt_compute_lrf", _I, [_P, c_int_p, _I, ctypes.POINTER(_SSA_LRF)]),
    ("ssa_opt_lrf_free", None, [ctypes.POINTER(_SSA_LRF)]),
    ("ssa_opt_forecast_with_lrf", _I, [ctypes.POINTER(_SSA_LRF), c_double_p, _I, _I, c_double_p]),
    ("ssa_opt_forecast", _I, [_P, c_int_p, _I, _I, c_double_p]),
    ("ssa_opt_forecast_full", _I, [_P, c_int_p, _I, _I, c_double_p]),
    ("ssa_opt_vforecast", _I, [_P, c_int_p, _I, _I, c_double_p]),
    ("ssa_opt_vforecast_full", _I, [_P, c_int_p, _I, _I, c_double_p]),
    ("ssa_opt_vforecast_fast", _I, [_P, c_int_p, _I, c_double_p, _I, _I, c_double_p]),
    ("ssa_opt_cadzow", _I, [c_double_p, _I, _I, _I, _I, _D, c_double_p, ctypes.POINTER(_SSA_CadzowResult)]),
    ("ssa_opt_cadzow_weighted", _I, [c_double_p, _I, _I, _I, _I, _D, _D, c_double_p,
                                     ctypes.POINTER(_SSA_CadzowResult)]),
    ("ssa_opt_cadzow_inplace", _I, [_P, _I, _I, _D, ctypes.POINTER(_SSA_CadzowResult)]),
    ("ssa_opt_gapfill", _I, [c_double_p, _I, _I, _I, _I, _D, ctypes.POINTER(_SSA_GapFillResult)]),
    ("ssa_opt_gapfill_simple", _I, [c_double_p, _I, _I, _I, ctypes.POINTER(_SSA_GapFillResult)]),
    ("ssa_opt_parestimate", _I, [_P, c_int_p, _I, ctypes.POINTER(_SSA_ParEstimate)]),
    ("ssa_opt_parestimate_free", None, [ctypes.POINTER(_SSA_ParEstimate)]),
    ("ssa_amd_esprit_poles", _I, [c_double_p, _I, _I, c_double_p, c_double_p]),
    ("mssa_opt_init", _I, [_M, c_double_p, _I, _I, _I]),
    ("mssa_opt_decompose", _I, [_M, _I, _I]),
    ("mssa_opt_reconstruct", _I, [_M, _I, c_int_p, _I, c_double_p]),
    ("mssa_opt_reconstruct_all", _I, [_M, c_int_p, _I, c_double_p]),
    ("mssa_opt_series_contributions", _I, [_M, c_double_p]),
    ("mssa_opt_variance_explained", _D, [_M, _I, _I]),
    ("mssa_opt_free", None, [_M]),
    ("ssa_amd_set_options", _I, [_P, _I, _I, _U64]),
    ("mssa_amd_set_options", _I, [_M, _I, _I, _U64]),
    ("ssa_amd_inject_omega", _I, [_P, c_double_p, _I, _I]),
    ("ssa_amd_reconstruct_many", _I, [_P, c_int_p, c_int_p, _I, c_double_p]),
    ("ssa_amd_plan_info", _I, [_P, c_int_p]),
    ("ssa_amd_push_samples", _I, [_P, c_double_p, _I]),
    ("ssa_amd_set_mirror", _I, [_P, _I]),
    ("ssa_amd_sync_mirrors", _I, [_P]),
    ("ssa_amd_batch_create", _VP, [_I, c_int_p, c_int_p, c_int_p, _I, _I, _I, _U64]),
    ("ssa_amd_batch_set_signals_f32", _I, [_VP, c_float_p, c_ll_p]),
    ("ssa_amd_batch_set_key", _I, [_VP, _I, _U64]),
    ("ssa_amd_batch_set_wide_min", _I, [_VP, _I]),
    ("ssa_amd_autotune", _I, [_I, c_double_p]),
    ("ssa_amd_batch_rebind", _I, [_VP, _I, c_int_p, c_int_p, c_int_p, c_int_p]),
    ("ssa_amd_batch_outputs", _I, [_VP, c_double_p, c_ll_p, c_int_p]),
    ("ssa_amd_batch_run", _I, [_VP, _I]),
    ("ssa_amd_batch_sync", _I, [_VP]),
    ("ssa_amd_batch_sigma", _I, [_VP, _I, c_double_p]),
    ("ssa_amd_batch_output", _I, [_VP, _I, c_double_p]),
    ("ssa_amd_batch_bytes", ctypes.c_longlong, [_VP]),
    ("ssa_amd_series_bytes", ctypes.c_longlong, [_I, _I, _I, _I, _I, _I, _I]),
    ("ssa_mkl_init", _I, [_I]),
    ("ssa_mkl_init_14900kf", _I, [_I]),
    ("ssa_mkl_init_generic", _I, [_I]),
    ("ssa_mkl_set_threads", None, [_I]),
    ("ssa_mkl_get_threads", _I, []),
    ("ssa_mkl_get_cpu_info", None, [c_int_p, c_int_p, c_int_p, c_int_p]),
    ("ssa_mkl_disable_affinity", None, []),
    ("ssa_amd_batch_free", None, [_VP]),
    ("ssa_amd_batch_profile", _I, [_VP, _I, c_int_p, c_int_p, c_float_p]),
    ("ssa_amd_batch_create_ex", _VP, [_I, c_int_p, c_int_p, c_int_p, c_int_p, _I, _I, _I, _I, _U64]),
    ("ssa_amd_batch_create_g", _VP, [_I, c_int_p, c_int_p, c_int_p, c_int_p, _I, _I, _I, _I, _U64, _I]),
    ("ssa_amd_batch_query", _I, [_VP]),
    ("ssa_amd_batch_status", _I, [_VP, _I]),
    ("ssa_amd_batch_plan", _I, [_VP, _I, c_int_p]),
    ("ssa_amd_reconstruct_noalloc", _I, [_VP, _VP, _VP, _I, _I, _I, c_int_p, _I, _VP, _VP]),
]


def load(path=LIB_PATH):
    if not os.path.exists(path):
        raise RuntimeError(f"libssa.so not found at {path}: build it (python -c 'import __graft_entry__ as g; g.build()')")
    lib = ctypes.CDLL(path)
    for name, res, args in SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def _dp(a):
    return a.ctypes.data_as(c_double_p)


def _ip(a):
    return a.ctypes.data_as(c_int_p)


def autotune(precision=1):
    """Measure the per-transform device cost at every FFT length (ns, for
    2^6..2^13) and plan every later context and batch of this precision with
    it (no counterpart in the reference)."""
    out = np.zeros(8)
    if lib().ssa_amd_autotune(int(precision), _dp(out)) != 0:
        raise RuntimeError("autotune failed")
    return {6 + i: float(v) for i, v in enumerate(out)}


def cadzow(x, L, rank, max_iter=20, tol=1e-9, alpha=1.0):
    """Cadzow iterations (MKL/py/ssa_wrapper.py:415-480): (signal, iterations,
    final_diff, converged); alpha < 1 blends with the input."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x))
    r = _SSA_CadzowResult()
    if alpha == 1.0:
        it = lib().ssa_opt_cadzow(_dp(x), len(x), L, rank, max_iter, tol, _dp(out), ctypes.byref(r))
    else:
        it = lib().ssa_opt_cadzow_weighted(_dp(x), len(x), L, rank, max_iter, tol, alpha, _dp(out), ctypes.byref(r))
    if it < 0:
        raise RuntimeError("Cadzow iterations failed")
    return out, r.iterations, r.final_diff, r.converged > 0.5


def series_bytes(N, L, k, p=10, q=2, precision=1, h=0):
    """Device bytes one series takes in a batch (host arithmetic; no device
    call): what a caller admits against a device-memory budget."""
    v = lib().ssa_amd_series_bytes(int(N), int(L), int(k), int(p), int(q), int(precision), int(h))
    if v < 0:
        raise ValueError("invalid series dimensions")
    return int(v)


def esprit_poles(U):
    """ESPRIT poles of the columns of U (L x k): the host core of
    ssa_opt_parestimate (no device call)."""
    U = np.asfortranarray(U, dtype=np.float64)
    L, k = U.shape
    wr, wi = np.zeros(k), np.zeros(k)
    if lib().ssa_amd_esprit_poles(U.ctypes.data_as(c_double_p), L, k, _dp(wr), _dp(wi)) != 0:
        raise RuntimeError("ESPRIT failed")
    return wr + 1j * wi


def gapfill(x, L, rank, max_iter=50, tol=1e-6, simple=False):
    """Fill the NaNs of a copy of x (MKL/ssa_opt.h:3498-3877): (filled,
    iterations, final_diff, converged, n_gaps)."""
    y = np.array(x, dtype=np.float64)
    r = _SSA_GapFillResult()
    if simple:
        rc = lib().ssa_opt_gapfill_simple(_dp(y), len(y), L, rank, ctypes.byref(r))
    else:
        rc = lib().ssa_opt_gapfill(_dp(y), len(y), L, rank, max_iter, tol, ctypes.byref(r))
    if rc != 0:
        raise RuntimeError("gap filling failed")
    return y, r.iterations, r.final_diff, bool(r.converged), r.n_gaps


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def _checked(rc, what):
    """Raise RuntimeError when a C call returned a failure code."""
    if rc != 0:
        raise RuntimeError(f"{what} (libssa returned {rc})")


# defaults of the wrapper's non-randomized methods (MKL/py/ssa_wrapper.py:704-745):
# block width min(k, BLOCK_CAP), MAX_ITER iterations
BLOCK_CAP = 32
MAX_ITER = 100


class _Device:
    """One C context (SSA_Opt or MSSA_Opt) owned by a Python object: the typed
    struct itself, so the library's own state (prepared capacity, component
    count, flags) is read from it rather than duplicated here."""

    _struct = None
    _free = None

    def __init__(self):
        self.lib = lib()
        self.ctx = self._struct()

    @property
    def ref(self):
        return ctypes.byref(self.ctx)

    def __del__(self):
        if getattr(self, "ctx", None) is not None and getattr(self, "lib", None) is not None:
            getattr(self.lib, self._free)(ctypes.byref(self.ctx))
            self.ctx = None

    def _require(self):
        if not self.ctx.decomposed:
            raise RuntimeError("no decomposition yet: call decompose() first")

    def _out(self, n, fn, *args, what):
        out = np.zeros(n)
        _checked(fn(self.ref, *args, _dp(out)), what)
        return out


class SSA(_Device):
    """Univariate SSA on the device, with the interface of the reference
    wrapper's SSA class (MKL/py/ssa_wrapper.py:600-1202): method names,
    arguments, ValueError for a series the library refuses, RuntimeError for a
    failed call.  q, precision and seed (no counterpart there) set the power
    iterations, the plain / mixed form and the test matrix's Philox key."""

    _struct = _SSA_Opt
    _free = "ssa_opt_free"

    def __init__(self, x, L: int, q: int = None, precision: int = None, seed: int = None):
        super().__init__()
        series = _f64(x)
        if self.lib.ssa_opt_init(self.ref, _dp(series), series.size, int(L)) != 0:
            raise ValueError(f"ssa_opt_init refused {series.size} samples with window {L} (needs 2 <= L <= N - 1)")
        if (q, precision, seed) != (None, None, None):
            self.set_options(2 if q is None else q, 0 if precision is None else precision,
                             42 if seed is None else seed)

    # dimensions and state, read from the context
    N = property(lambda self: self.ctx.N)
    L = property(lambda self: self.ctx.L)
    K = property(lambda self: self.ctx.K)
    n_components = property(lambda self: self.ctx.n_components if self.ctx.decomposed else 0)

    # --- extensions --------------------------------------------------------
    def set_options(self, q=2, precision=0, seed=42):
        _checked(self.lib.ssa_amd_set_options(self.ref, int(q), int(precision), int(seed)),
                 f"options q={q} precision={precision} refused")
        return self

    def inject_omega(self, omega):
        om = np.asfortranarray(omega, dtype=np.float64)
        _checked(self.lib.ssa_amd_inject_omega(self.ref, _dp(om), om.shape[0], om.shape[1]),
                 f"test matrix of shape {om.shape} refused")
        return self

    def plan_info(self):
        out = np.zeros(6, dtype=np.int32)
        self.lib.ssa_amd_plan_info(self.ref, _ip(out))
        return dict(zip(("n", "CL", "CK", "AL", "SK", "sweeps"), (int(v) for v in out)))

    def reconstruct_many(self, groups: List[List[int]]):
        flat = _i32([i for g in groups for i in g])
        sizes = _i32([len(g) for g in groups])
        self._require()
        out = np.zeros(len(groups) * self.N)
        _checked(self.lib.ssa_amd_reconstruct_many(self.ref, _ip(flat), _ip(sizes), len(groups), _dp(out)),
                 f"reconstruction of {len(groups)} groups failed")
        return out.reshape(len(groups), self.N)

    def hankel_matvec(self, v):
        y = np.zeros(self.L)
        self.lib.ssa_opt_hankel_matvec(self.ref, _dp(_f64(v)), _dp(y))
        return y

    def hankel_matvec_T(self, u):
        z = np.zeros(self.K)
        self.lib.ssa_opt_hankel_matvec_T(self.ref, _dp(_f64(u)), _dp(z))
        return z

    def set_mirror(self, on: bool) -> "SSA":
        """Host U/V mirrors refreshed by every decomposition (on) or on demand (off)."""
        _checked(self.lib.ssa_amd_set_mirror(self.ref, int(bool(on))), "set_mirror refused")
        return self

    def push(self, samples) -> "SSA":
        """Slide the window by len(samples) new samples (ssa_amd_push_samples)."""
        new = _f64(samples)
        _checked(self.lib.ssa_amd_push_samples(self.ref, _dp(new), new.size),
                 f"push of {new.size} samples refused (1 <= m <= N)")
        return self

    def factors(self):
        """(U, V, sigma) copies of the current decomposition: U is L x k and V
        K x k (column-major host mirrors, refreshed first when stale)."""
        self._require()
        _checked(self.lib.ssa_amd_sync_mirrors(self.ref), "mirror refresh failed")
        k = self.ctx.n_components
        U = np.ctypeslib.as_array(self.ctx.U, shape=(k, self.L)).T.copy()
        V = np.ctypeslib.as_array(self.ctx.V, shape=(k, self.K)).T.copy()
        return U, V, np.ctypeslib.as_array(self.ctx.sigma, shape=(k,)).copy()

    def cache_ffts(self) -> "SSA":
        """Keep the component spectra on the device (ssa_opt_cache_ffts) until
        the next decomposition or update."""
        self._require()
        _checked(self.lib.ssa_opt_cache_ffts(self.ref), "component spectra cache failed")
        return self

    def free_cached_ffts(self) -> "SSA":
        self.lib.ssa_opt_free_cached_ffts(self.ref)
        return self

    fft_cached = property(lambda self: bool(self.ctx.fft_cached))

    def wcorr_matrix_fast(self) -> np.ndarray:
        """W-correlation from the cached spectra (the slow path without a cache)."""
        self._require()
        k = self.ctx.n_components
        return self._out(k * k, self.lib.ssa_opt_wcorr_matrix_fast, what="W-correlation failed").reshape(k, k)

    # --- the reference wrapper's interface ------------------------------------
    def prepare(self, max_k: int, oversampling: int = 8) -> "SSA":
        _checked(self.lib.ssa_opt_prepare(self.ref, int(max_k), int(oversampling)),
                 f"prepare for {max_k} components + {oversampling} refused (k + p <= {BMAX})")
        return self

    def update_signal(self, new_x) -> "SSA":
        series = _f64(new_x)
        if series.size != self.N:
            raise ValueError(f"update_signal needs {self.N} samples, got {series.size}")
        _checked(self.lib.ssa_opt_update_signal(self.ref, _dp(series)), "update_signal failed")
        return self

    def _randomized(self, k, opts):
        p = opts.get("oversampling", 8)
        width = min(k + (p if p > 0 else 8), self.L, self.K)
        # prepare on demand, as the wrapper does, from the context's own capacity
        if not self.ctx.prepared or width > self.ctx.prepared_kp:
            self.prepare(k, p)
        return self.lib.ssa_opt_decompose_randomized(self.ref, k, p)

    def _block(self, k, opts):
        return self.lib.ssa_opt_decompose_block(self.ref, k, opts.get("block_size", min(k, BLOCK_CAP)),
                                                opts.get("max_iter", MAX_ITER))

    def _sequential(self, k, opts):
        return self.lib.ssa_opt_decompose(self.ref, k, opts.get("max_iter", MAX_ITER))

    _METHODS = {"randomized": _randomized, "block": _block, "sequential": _sequential}

    def decompose(self, k: int, method: str = "randomized", **kwargs) -> "SSA":
        run = self._METHODS.get(method)
        if run is None:
            raise ValueError(f"method {method!r} is not one of {sorted(self._METHODS)}")
        _checked(run(self, int(k), kwargs), f"{method} decomposition into {k} components failed")
        return self

    @property
    def singular_values(self) -> np.ndarray:
        self._require()
        return self._getter(self.lib.ssa_opt_get_singular_values)

    @property
    def eigenvalues(self) -> np.ndarray:
        self._require()
        return self._getter(self.lib.ssa_opt_get_eigenvalues)

    def _getter(self, fn):
        out = np.zeros(self.ctx.n_components)
        n = fn(self.ref, _dp(out), out.size)
        return out[:max(n, 0)]

    @property
    def total_variance(self) -> float:
        self._require()
        return self.lib.ssa_opt_get_total_variance(self.ref)

    def _grouped(self, fn, group, n, *extra, what):
        self._require()
        g = _i32(group)
        return self._out(n, fn, _ip(g), g.size, *extra, what=what)

    def reconstruct(self, group: List[int]) -> np.ndarray:
        return self._grouped(self.lib.ssa_opt_reconstruct, group, self.N, what=f"reconstruction of {list(group)} failed")

    def forecast(self, group: List[int], n_forecast: int) -> np.ndarray:
        return self._grouped(self.lib.ssa_opt_forecast, group, n_forecast, int(n_forecast),
                             what="R-forecast failed (a group index out of range, or verticality >= 1)")

    def forecast_full(self, group: List[int], n_forecast: int) -> np.ndarray:
        return self._grouped(self.lib.ssa_opt_forecast_full, group, self.N + n_forecast, int(n_forecast),
                             what="R-forecast failed (a group index out of range, or verticality >= 1)")

    def vforecast(self, group: List[int], n_forecast: int) -> np.ndarray:
        """V-forecast (MKL/py/ssa_wrapper.py:880-900, MKL/ssa_opt.h:2906-2965)."""
        return self._grouped(self.lib.ssa_opt_vforecast, group, n_forecast, int(n_forecast),
                             what="V-forecast failed (a group index out of range, or verticality >= 1)")

    def vforecast_full(self, group: List[int], n_forecast: int) -> np.ndarray:
        return self._grouped(self.lib.ssa_opt_vforecast_full, group, self.N + n_forecast, int(n_forecast),
                             what="V-forecast failed (a group index out of range, or verticality >= 1)")

    def vforecast_fast(self, group: List[int], base_signal, n_forecast: int) -> np.ndarray:
        """V-forecast continuing base_signal (its last L-1 values)."""
        base = _f64(base_signal)
        return self._grouped(self.lib.ssa_opt_vforecast_fast, group, n_forecast, _dp(base), base.size,
                             int(n_forecast), what="V-forecast failed (base shorter than L - 1, or bad group)")

    def cadzow_inplace(self, rank: int, max_iter: int = 20, tol: float = 1e-9):
        """Cadzow on this context's signal (MKL/ssa_opt.h:3319-3440):
        (iterations, final_diff, converged); the context keeps the result."""
        r = _SSA_CadzowResult()
        if self.lib.ssa_opt_cadzow_inplace(self.ref, int(rank), int(max_iter), float(tol), ctypes.byref(r)) < 0:
            raise RuntimeError(f"Cadzow iterations at rank {rank} failed")
        return r.iterations, r.final_diff, r.converged > 0.5

    def parestimate(self, group: List[int] = None) -> dict:
        """ESPRIT periods, frequencies, moduli and rates of the group's
        eigenvectors (all computed components when group is None;
        MKL/ssa_opt.h:4304-4558)."""
        self._require()
        r = _SSA_ParEstimate()
        g = None if group is None else _i32(group)
        rc = self.lib.ssa_opt_parestimate(self.ref, None if g is None else _ip(g), 0 if g is None else g.size,
                                          ctypes.byref(r))
        _checked(rc, "ESPRIT estimation failed")
        fields = ("periods", "frequencies", "moduli", "rates")
        est = {f: np.ctypeslib.as_array(getattr(r, f), shape=(r.n_components,)).copy() for f in fields}
        self.lib.ssa_opt_parestimate_free(ctypes.byref(r))
        return est

    def lrf(self, group: List[int]):
        """(status, R coefficients or None, verticality) of ssa_opt_compute_lrf."""
        self._require()
        g = _i32(group)
        f = _SSA_LRF()
        rc = self.lib.ssa_opt_compute_lrf(self.ref, _ip(g), g.size, ctypes.byref(f))
        coef = np.ctypeslib.as_array(f.R, shape=(self.L - 1,)).copy() if rc == 0 else None
        nu = f.verticality
        self.lib.ssa_opt_lrf_free(ctypes.byref(f))
        return rc, coef, nu

    def get_trend(self) -> np.ndarray:
        self._require()
        return self._out(self.N, self.lib.ssa_opt_get_trend, what="trend reconstruction failed")

    def get_noise(self, noise_start: int) -> np.ndarray:
        self._require()
        return self._out(self.N, self.lib.ssa_opt_get_noise, int(noise_start), what="noise reconstruction failed")

    def wcorr_matrix(self) -> np.ndarray:
        self._require()
        k = self.ctx.n_components
        return self._out(k * k, self.lib.ssa_opt_wcorr_matrix, what="W-correlation failed").reshape(k, k)

    def wcorr(self, i: int, j: int) -> float:
        self._require()
        return self.lib.ssa_opt_wcorr_pair(self.ref, int(i), int(j))

    def variance_explained(self, start: int = 0, end: int = -1) -> float:
        self._require()
        return self.lib.ssa_opt_variance_explained(self.ref, int(start), int(end))

    def find_periodic_pairs(self, max_pairs: int = 10, sv_tol: float = 0.1, wcorr_thresh: float = 0.5):
        self._require()
        pairs = np.zeros(2 * max_pairs, dtype=np.int32)
        n = self.lib.ssa_opt_find_periodic_pairs(self.ref, _ip(pairs), int(max_pairs), float(sv_tol),
                                                 float(wcorr_thresh))
        return [tuple(int(v) for v in pairs[2 * i:2 * i + 2]) for i in range(n)]


class MSSA(_Device):
    """Multivariate SSA on the device, with the interface of the reference
    wrapper's MSSA class (MKL/py/ssa_wrapper.py:1208-1371)."""

    _struct = _MSSA_Opt
    _free = "mssa_opt_free"

    def __init__(self, X, L: int, q: int = None, precision: int = None, seed: int = None):
        super().__init__()
        series = _f64(X)
        if series.ndim != 2:
            raise ValueError(f"MSSA takes an (M, N) array of channels, got {series.ndim} dimensions")
        M, N = series.shape
        if self.lib.mssa_opt_init(self.ref, _dp(series), M, N, int(L)) != 0:
            raise ValueError(f"mssa_opt_init refused {M} channels of {N} samples with window {L}")
        if (q, precision, seed) != (None, None, None):
            _checked(self.lib.mssa_amd_set_options(self.ref, 2 if q is None else q, 0 if precision is None else precision,
                                                   42 if seed is None else seed), "MSSA options refused")

    M = property(lambda self: self.ctx.M)
    N = property(lambda self: self.ctx.N)
    L = property(lambda self: self.ctx.L)
    K = property(lambda self: self.ctx.K)
    n_components = property(lambda self: self.ctx.n_components if self.ctx.decomposed else 0)

    def decompose(self, k: int, oversampling: int = 8) -> "MSSA":
        _checked(self.lib.mssa_opt_decompose(self.ref, int(k), int(oversampling)),
                 f"MSSA decomposition into {k} components failed")
        return self

    def reconstruct(self, series_idx: int, group: List[int]) -> np.ndarray:
        self._require()
        g = _i32(group)
        return self._out(self.N, self.lib.mssa_opt_reconstruct, int(series_idx), _ip(g), g.size,
                         what=f"reconstruction of channel {series_idx} failed")

    def reconstruct_all(self, group: List[int]) -> np.ndarray:
        self._require()
        g = _i32(group)
        out = self._out(self.M * self.N, self.lib.mssa_opt_reconstruct_all, _ip(g), g.size,
                        what="reconstruction of every channel failed")
        return out.reshape(self.M, self.N)

    def series_contributions(self) -> np.ndarray:
        self._require()
        k = self.ctx.n_components
        return self._out(self.M * k, self.lib.mssa_opt_series_contributions,
                         what="channel contributions failed").reshape(self.M, k)

    def variance_explained(self, start: int = 0, end: int = -1) -> float:
        return self.lib.mssa_opt_variance_explained(self.ref, int(start), int(end))


class Batch:
    """Many uneven series in one device engine (no counterpart in the reference)."""

    def __init__(self, xs, Ls, ks, p=10, q=2, precision=1, seed=4, groups=None, h=0, keys=None, branches=None):
        """groups[i]: reconstruct (and forecast) components 0..groups[i]-1
        of series i (default all k); h > 0 appends an h-step LRF forecast;
        keys[i]: Philox key of series i (default derived from seed and i);
        branches: parallel graph branches (1..4; default SSA_AMD_BRANCHES or 2)."""
        self._l = lib()
        self.horizon = int(h)
        self._dims(xs, Ls, ks)
        g = None if groups is None else np.ascontiguousarray(groups, dtype=np.int32)
        self._h = self._l.ssa_amd_batch_create_g(self.count, _ip(self.N), _ip(self.L), _ip(self.k),
                                                None if g is None else _ip(g), self.horizon, p, q, precision, seed,
                                                int(branches or 0))
        if not self._h:
            raise RuntimeError("batch creation failed")
        self._upload(xs, keys)

    def _dims(self, xs, Ls, ks):
        self.count = len(xs)
        self.N = np.ascontiguousarray([len(x) for x in xs], dtype=np.int32)
        self.L = np.ascontiguousarray(Ls, dtype=np.int32)
        self.k = np.ascontiguousarray(ks, dtype=np.int32)

    def _upload(self, xs, keys):
        cat = np.ascontiguousarray(np.concatenate([np.asarray(x, dtype=np.float32) for x in xs]))
        offs = np.zeros(self.count, dtype=np.int64)
        offs[1:] = np.cumsum(self.N[:-1])
        if self._l.ssa_amd_batch_set_signals_f32(self._h, cat.ctypes.data_as(c_float_p),
                                                 offs.ctypes.data_as(c_ll_p)) != 0:
            raise RuntimeError("batch upload failed")
        if keys is not None:
            for i, key in enumerate(keys):
                if self._l.ssa_amd_batch_set_key(self._h, i, int(key) & ((1 << 64) - 1)) != 0:
                    raise RuntimeError("batch key failed")

    def set_wide_min(self, n):
        """Block-matvec launches of at least n items take the wide transform
        geometry (1: every launch), for this batch and every rebind of it."""
        if self._l.ssa_amd_batch_set_wide_min(self._h, int(n)) != 0:
            raise RuntimeError("batch geometry setting failed")
        return self

    def rebind(self, xs, Ls, ks, groups=None, keys=None):
        """Re-target this (idle) batch to new series with the same p, q, h,
        precision and seed; device buffers are reused when large enough."""
        self._dims(xs, Ls, ks)
        g = None if groups is None else np.ascontiguousarray(groups, dtype=np.int32)
        if self._l.ssa_amd_batch_rebind(self._h, self.count, _ip(self.N), _ip(self.L), _ip(self.k),
                                        None if g is None else _ip(g)) != 0:
            raise RuntimeError("batch rebind failed")
        self._upload(xs, keys)

    def __del__(self):
        if getattr(self, "_h", None):
            self._l.ssa_amd_batch_free(self._h)
            self._h = None

    def run(self, use_graph=True):
        if self._l.ssa_amd_batch_run(self._h, int(use_graph)) != 0:
            raise RuntimeError("batch run failed")

    def sync(self):
        if self._l.ssa_amd_batch_sync(self._h) != 0:
            raise RuntimeError("batch sync failed")

    def sigma(self, i):
        out = np.zeros(int(self.k[i]) + 80)
        n = self._l.ssa_amd_batch_sigma(self._h, i, _dp(out))
        if n < 0:
            raise RuntimeError("batch sigma failed")
        return out[:n]

    def query(self):
        r = self._l.ssa_amd_batch_query(self._h)
        if r < 0:
            raise RuntimeError("batch query failed")
        return bool(r)

    def status(self, i):
        return int(self._l.ssa_amd_batch_status(self._h, i))

    def output(self, i):
        out = np.zeros(int(self.N[i]) + self.horizon)
        if self._l.ssa_amd_batch_output(self._h, i, _dp(out)) != 0:
            raise RuntimeError("batch output failed")
        return out

    def outputs(self):
        """(list of every series' output, status words) with one device
        synchronisation per branch."""
        lens = self.N.astype(np.int64) + self.horizon
        offs = np.zeros(self.count, dtype=np.int64)
        offs[1:] = np.cumsum(lens[:-1])
        cat = np.zeros(int(lens.sum()))
        st = np.zeros(self.count, dtype=np.int32)
        if self._l.ssa_amd_batch_outputs(self._h, _dp(cat), offs.ctypes.data_as(c_ll_p), _ip(st)) != 0:
            raise RuntimeError("batch outputs failed")
        return [cat[offs[i]:offs[i] + lens[i]] for i in range(self.count)], st

    def bytes(self):
        return int(self._l.ssa_amd_batch_bytes(self._h))

    PHASES = ("spectra", "matvec_H", "matvec_HT", "gram", "chol", "apply", "svd", "uv", "recon", "olap", "forecast",
              "omega")

    def profile(self, max_rec=4096):
        """One eager step with HIP events around every launch:
        list of (phase name, log2 n or panel id, ms)."""
        ph = np.zeros(max_rec, dtype=np.int32)
        lg = np.zeros(max_rec, dtype=np.int32)
        ms = np.zeros(max_rec, dtype=np.float32)
        n = self._l.ssa_amd_batch_profile(self._h, max_rec, _ip(ph), _ip(lg), ms.ctypes.data_as(c_float_p))
        if n < 0:
            raise RuntimeError("batch profile failed")
        return [(self.PHASES[ph[i]], int(lg[i]), float(ms[i])) for i in range(n)]

    def plan(self, i):
        out = np.zeros(8, dtype=np.int32)
        if self._l.ssa_amd_batch_plan(self._h, i, _ip(out)) != 0:
            raise RuntimeError("batch plan failed")
        return dict(zip(("n", "CL", "CK", "AL", "SK", "b", "k", "lgn"), (int(v) for v in out)))
