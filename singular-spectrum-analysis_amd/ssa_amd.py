"""Python front end of the device SSA library (libssa.so), mirroring the
reference's ctypes wrapper MKL/py/ssa_wrapper.py: the same classes (SSA,
MSSA), method names, argument meaning and error behaviour (ValueError on a
failed init, RuntimeError on a failed call), on top of the same C symbols.

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


class _SSA_Opt(ctypes.Structure):
    # opaque block, as MKL/py/ssa_wrapper.py:92-94
    _fields_ = [("_opaque", ctypes.c_char * 1024)]


class _MSSA_Opt(ctypes.Structure):
    _fields_ = [("_opaque", ctypes.c_char * 1024)]


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


class SSA:
    """Univariate SSA on the device (MKL/py/ssa_wrapper.py:600-1202)."""

    def __init__(self, x, L: int, q: int = None, precision: int = None, seed: int = None):
        self._l = lib()
        self._ctx = _SSA_Opt()
        self._x = np.ascontiguousarray(x, dtype=np.float64)
        self.N = len(self._x)
        self.L = L
        self.K = self.N - L + 1
        self.n_components = 0
        self._decomposed = False
        self._prepared = False
        self._prepared_k = 0
        self._prepared_oversampling = 0
        if self._l.ssa_opt_init(ctypes.byref(self._ctx), _dp(self._x), self.N, L) != 0:
            raise ValueError(f"SSA initialization failed (N={self.N}, L={L})")
        if q is not None or precision is not None or seed is not None:
            self.set_options(2 if q is None else q, 0 if precision is None else precision,
                             42 if seed is None else seed)

    def __del__(self):
        if hasattr(self, "_ctx") and getattr(self, "_l", None) is not None:
            self._l.ssa_opt_free(ctypes.byref(self._ctx))

    # --- extensions -------------------------------------------------------
    def set_options(self, q=2, precision=0, seed=42):
        if self._l.ssa_amd_set_options(ctypes.byref(self._ctx), q, precision, seed) != 0:
            raise RuntimeError("set_options failed")
        self._prepared = False
        self._decomposed = False
        return self

    def inject_omega(self, omega):
        om = np.asfortranarray(omega, dtype=np.float64)
        if self._l.ssa_amd_inject_omega(ctypes.byref(self._ctx), _dp(om), om.shape[0], om.shape[1]) != 0:
            raise RuntimeError("inject_omega failed")
        return self

    def plan_info(self):
        out = np.zeros(6, dtype=np.int32)
        self._l.ssa_amd_plan_info(ctypes.byref(self._ctx), _ip(out))
        return dict(zip(("n", "CL", "CK", "AL", "SK", "sweeps"), (int(v) for v in out)))

    def reconstruct_many(self, groups: List[List[int]]):
        gl = np.ascontiguousarray([i for g in groups for i in g], dtype=np.int32)
        gc = np.ascontiguousarray([len(g) for g in groups], dtype=np.int32)
        out = np.zeros(len(groups) * self.N)
        if self._l.ssa_amd_reconstruct_many(ctypes.byref(self._ctx), _ip(gl), _ip(gc), len(groups), _dp(out)) != 0:
            raise RuntimeError("reconstruct_many failed")
        return out.reshape(len(groups), self.N)

    def hankel_matvec(self, v):
        v = np.ascontiguousarray(v, dtype=np.float64)
        y = np.zeros(self.L)
        self._l.ssa_opt_hankel_matvec(ctypes.byref(self._ctx), _dp(v), _dp(y))
        return y

    def hankel_matvec_T(self, u):
        u = np.ascontiguousarray(u, dtype=np.float64)
        z = np.zeros(self.K)
        self._l.ssa_opt_hankel_matvec_T(ctypes.byref(self._ctx), _dp(u), _dp(z))
        return z

    # --- the reference wrapper's interface ----------------------------------
    def prepare(self, max_k: int, oversampling: int = 8) -> "SSA":
        if self._l.ssa_opt_prepare(ctypes.byref(self._ctx), max_k, oversampling) != 0:
            raise RuntimeError(f"prepare() failed (max_k={max_k})")
        self._prepared = True
        self._prepared_k = max_k
        self._prepared_oversampling = oversampling
        return self

    def update_signal(self, new_x) -> "SSA":
        new_x = np.ascontiguousarray(new_x, dtype=np.float64)
        if len(new_x) != self.N:
            raise ValueError(f"Signal length mismatch: {len(new_x)} vs {self.N}")
        if self._l.ssa_opt_update_signal(ctypes.byref(self._ctx), _dp(new_x)) != 0:
            raise RuntimeError("update_signal() failed")
        self._x = new_x
        return self

    def set_mirror(self, on: bool) -> "SSA":
        """Host U/V mirrors refreshed by every decomposition (on) or on demand (off)."""
        if self._l.ssa_amd_set_mirror(ctypes.byref(self._ctx), int(bool(on))) != 0:
            raise RuntimeError("set_mirror failed")
        return self

    def push(self, samples) -> "SSA":
        """Slide the window by len(samples) new samples (ssa_amd_push_samples)."""
        a = np.ascontiguousarray(samples, dtype=np.float64)
        if self._l.ssa_amd_push_samples(ctypes.byref(self._ctx), _dp(a), len(a)) != 0:
            raise RuntimeError("push_samples failed")
        self._x = np.concatenate([self._x[len(a):], a]) if len(a) < self.N else a[-self.N:].copy()
        self._decomposed = False
        return self

    def decompose(self, k: int, method: str = "randomized", **kwargs) -> "SSA":
        if method == "randomized":
            oversampling = kwargs.get("oversampling", 8)
            kp = k + oversampling
            if not self._prepared or kp > self._prepared_k + self._prepared_oversampling:
                self.prepare(k, oversampling)
            ret = self._l.ssa_opt_decompose_randomized(ctypes.byref(self._ctx), k, oversampling)
        elif method == "block":
            ret = self._l.ssa_opt_decompose_block(ctypes.byref(self._ctx), k, kwargs.get("block_size", min(k, 32)),
                                                  kwargs.get("max_iter", 100))
        elif method == "sequential":
            ret = self._l.ssa_opt_decompose(ctypes.byref(self._ctx), k, kwargs.get("max_iter", 100))
        else:
            raise ValueError(f"Unknown method: {method}")
        if ret != 0:
            raise RuntimeError(f"Decomposition failed (method={method}, k={k})")
        self.n_components = k
        self._decomposed = True
        return self

    def _check(self):
        if not self._decomposed:
            raise RuntimeError("Call decompose() first")

    @property
    def singular_values(self) -> np.ndarray:
        self._check()
        out = np.zeros(self.n_components)
        n = self._l.ssa_opt_get_singular_values(ctypes.byref(self._ctx), _dp(out), self.n_components)
        return out[:n]

    @property
    def eigenvalues(self) -> np.ndarray:
        self._check()
        out = np.zeros(self.n_components)
        n = self._l.ssa_opt_get_eigenvalues(ctypes.byref(self._ctx), _dp(out), self.n_components)
        return out[:n]

    @property
    def total_variance(self) -> float:
        self._check()
        return self._l.ssa_opt_get_total_variance(ctypes.byref(self._ctx))

    def reconstruct(self, group: List[int]) -> np.ndarray:
        self._check()
        out = np.zeros(self.N)
        g = np.ascontiguousarray(group, dtype=np.int32)
        if self._l.ssa_opt_reconstruct(ctypes.byref(self._ctx), _ip(g), len(group), _dp(out)) != 0:
            raise RuntimeError("Reconstruction failed")
        return out

    def forecast(self, group: List[int], n_forecast: int) -> np.ndarray:
        self._check()
        out = np.zeros(n_forecast)
        g = np.ascontiguousarray(group, dtype=np.int32)
        if self._l.ssa_opt_forecast(ctypes.byref(self._ctx), _ip(g), len(group), n_forecast, _dp(out)) != 0:
            raise RuntimeError("Forecast failed (verticality >= 1?)")
        return out

    def forecast_full(self, group: List[int], n_forecast: int) -> np.ndarray:
        self._check()
        out = np.zeros(self.N + n_forecast)
        g = np.ascontiguousarray(group, dtype=np.int32)
        if self._l.ssa_opt_forecast_full(ctypes.byref(self._ctx), _ip(g), len(group), n_forecast, _dp(out)) != 0:
            raise RuntimeError("Forecast failed")
        return out

    def vforecast(self, group: List[int], n_forecast: int) -> np.ndarray:
        """V-forecast (MKL/py/ssa_wrapper.py:880-900, MKL/ssa_opt.h:2906-2965)."""
        self._check()
        out = np.zeros(n_forecast)
        g = np.ascontiguousarray(group, dtype=np.int32)
        if self._l.ssa_opt_vforecast(ctypes.byref(self._ctx), _ip(g), len(group), n_forecast, _dp(out)) != 0:
            raise RuntimeError("V-forecast failed")
        return out

    def vforecast_full(self, group: List[int], n_forecast: int) -> np.ndarray:
        self._check()
        out = np.zeros(self.N + n_forecast)
        g = np.ascontiguousarray(group, dtype=np.int32)
        if self._l.ssa_opt_vforecast_full(ctypes.byref(self._ctx), _ip(g), len(group), n_forecast, _dp(out)) != 0:
            raise RuntimeError("V-forecast failed")
        return out

    def vforecast_fast(self, group: List[int], base_signal, n_forecast: int) -> np.ndarray:
        """V-forecast continuing base_signal (its last L-1 values)."""
        self._check()
        out = np.zeros(n_forecast)
        g = np.ascontiguousarray(group, dtype=np.int32)
        base = np.ascontiguousarray(base_signal, dtype=np.float64)
        if self._l.ssa_opt_vforecast_fast(ctypes.byref(self._ctx), _ip(g), len(group), _dp(base), len(base),
                                          n_forecast, _dp(out)) != 0:
            raise RuntimeError("V-forecast failed")
        return out

    def cadzow_inplace(self, rank: int, max_iter: int = 20, tol: float = 1e-9):
        """Cadzow on this context's signal (MKL/ssa_opt.h:3319-3440):
        (iterations, final_diff, converged); the context keeps the result."""
        r = _SSA_CadzowResult()
        it = self._l.ssa_opt_cadzow_inplace(ctypes.byref(self._ctx), rank, max_iter, tol, ctypes.byref(r))
        if it < 0:
            raise RuntimeError("Cadzow iterations failed")
        self.n_components = rank
        self._decomposed = True
        return r.iterations, r.final_diff, r.converged > 0.5

    def parestimate(self, group: List[int] = None) -> dict:
        """ESPRIT periods, frequencies, moduli and rates of the group's
        eigenvectors (all computed components when group is None;
        MKL/ssa_opt.h:4304-4558)."""
        self._check()
        r = _SSA_ParEstimate()
        if group is None:
            rc = self._l.ssa_opt_parestimate(ctypes.byref(self._ctx), None, 0, ctypes.byref(r))
        else:
            g = np.ascontiguousarray(group, dtype=np.int32)
            rc = self._l.ssa_opt_parestimate(ctypes.byref(self._ctx), _ip(g), len(g), ctypes.byref(r))
        if rc != 0:
            raise RuntimeError("ESPRIT failed")
        n = r.n_components
        out = {f: np.ctypeslib.as_array(getattr(r, f), shape=(n,)).copy()
               for f in ("periods", "frequencies", "moduli", "rates")}
        self._l.ssa_opt_parestimate_free(ctypes.byref(r))
        return out

    def lrf(self, group: List[int]):
        self._check()
        g = np.ascontiguousarray(group, dtype=np.int32)
        lrf = _SSA_LRF()
        rc = self._l.ssa_opt_compute_lrf(ctypes.byref(self._ctx), _ip(g), len(group), ctypes.byref(lrf))
        R = np.ctypeslib.as_array(lrf.R, shape=(self.L - 1,)).copy() if rc == 0 else None
        nu = lrf.verticality
        self._l.ssa_opt_lrf_free(ctypes.byref(lrf))
        return rc, R, nu

    def get_trend(self) -> np.ndarray:
        self._check()
        out = np.zeros(self.N)
        if self._l.ssa_opt_get_trend(ctypes.byref(self._ctx), _dp(out)) != 0:
            raise RuntimeError("get_trend failed")
        return out

    def get_noise(self, noise_start: int) -> np.ndarray:
        self._check()
        out = np.zeros(self.N)
        if self._l.ssa_opt_get_noise(ctypes.byref(self._ctx), noise_start, _dp(out)) != 0:
            raise RuntimeError("get_noise failed")
        return out

    def wcorr_matrix(self) -> np.ndarray:
        self._check()
        W = np.zeros((self.n_components, self.n_components))
        if self._l.ssa_opt_wcorr_matrix(ctypes.byref(self._ctx), _dp(W)) != 0:
            raise RuntimeError("W-correlation computation failed")
        return W

    def wcorr(self, i: int, j: int) -> float:
        self._check()
        return self._l.ssa_opt_wcorr_pair(ctypes.byref(self._ctx), i, j)

    def variance_explained(self, start: int = 0, end: int = -1) -> float:
        self._check()
        return self._l.ssa_opt_variance_explained(ctypes.byref(self._ctx), start, end)

    def find_periodic_pairs(self, max_pairs: int = 10, sv_tol: float = 0.1, wcorr_thresh: float = 0.5):
        self._check()
        pairs = np.zeros(2 * max_pairs, dtype=np.int32)
        n = self._l.ssa_opt_find_periodic_pairs(ctypes.byref(self._ctx), _ip(pairs), max_pairs, sv_tol, wcorr_thresh)
        return [(int(pairs[2 * i]), int(pairs[2 * i + 1])) for i in range(n)]


class MSSA:
    """Multivariate SSA on the device (MKL/py/ssa_wrapper.py:1208-1371)."""

    def __init__(self, X, L: int, q: int = None, precision: int = None, seed: int = None):
        self._l = lib()
        self._ctx = _MSSA_Opt()
        self._X = np.ascontiguousarray(X, dtype=np.float64)
        if self._X.ndim != 2:
            raise ValueError("X must be 2D array (M, N)")
        self.M, self.N = self._X.shape
        self.L = L
        self.K = self.N - L + 1
        self.n_components = 0
        self._decomposed = False
        if self._l.mssa_opt_init(ctypes.byref(self._ctx), _dp(self._X), self.M, self.N, L) != 0:
            raise ValueError(f"MSSA initialization failed (M={self.M}, N={self.N}, L={L})")
        if q is not None or precision is not None or seed is not None:
            if self._l.mssa_amd_set_options(ctypes.byref(self._ctx), 2 if q is None else q,
                                            0 if precision is None else precision,
                                            42 if seed is None else seed) != 0:
                raise RuntimeError("set_options failed")

    def __del__(self):
        if hasattr(self, "_ctx") and getattr(self, "_l", None) is not None:
            self._l.mssa_opt_free(ctypes.byref(self._ctx))

    def decompose(self, k: int, oversampling: int = 8) -> "MSSA":
        if self._l.mssa_opt_decompose(ctypes.byref(self._ctx), k, oversampling) != 0:
            raise RuntimeError(f"MSSA decomposition failed (k={k})")
        self.n_components = k
        self._decomposed = True
        return self

    def reconstruct(self, series_idx: int, group: List[int]) -> np.ndarray:
        if not self._decomposed:
            raise RuntimeError("Call decompose() first")
        out = np.zeros(self.N)
        g = np.ascontiguousarray(group, dtype=np.int32)
        if self._l.mssa_opt_reconstruct(ctypes.byref(self._ctx), series_idx, _ip(g), len(group), _dp(out)) != 0:
            raise RuntimeError("MSSA reconstruction failed")
        return out

    def reconstruct_all(self, group: List[int]) -> np.ndarray:
        if not self._decomposed:
            raise RuntimeError("Call decompose() first")
        out = np.zeros(self.M * self.N)
        g = np.ascontiguousarray(group, dtype=np.int32)
        if self._l.mssa_opt_reconstruct_all(ctypes.byref(self._ctx), _ip(g), len(group), _dp(out)) != 0:
            raise RuntimeError("MSSA reconstruction failed")
        return out.reshape(self.M, self.N)

    def series_contributions(self) -> np.ndarray:
        if not self._decomposed:
            raise RuntimeError("Call decompose() first")
        out = np.zeros(self.M * self.n_components)
        if self._l.mssa_opt_series_contributions(ctypes.byref(self._ctx), _dp(out)) != 0:
            raise RuntimeError("series_contributions failed")
        return out.reshape(self.M, self.n_components)

    def variance_explained(self, start: int = 0, end: int = -1) -> float:
        return self._l.mssa_opt_variance_explained(ctypes.byref(self._ctx), start, end)


class Batch:
    """Many uneven series in one device engine (no counterpart in the reference)."""

    def __init__(self, xs, Ls, ks, p=10, q=2, precision=1, seed=4, groups=None, h=0, keys=None):
        """groups[i]: reconstruct (and forecast) components 0..groups[i]-1
        of series i (default all k); h > 0 appends an h-step LRF forecast;
        keys[i]: Philox key of series i (default derived from seed and i)."""
        self._l = lib()
        self.horizon = int(h)
        self._dims(xs, Ls, ks)
        g = None if groups is None else np.ascontiguousarray(groups, dtype=np.int32)
        self._h = self._l.ssa_amd_batch_create_ex(self.count, _ip(self.N), _ip(self.L), _ip(self.k),
                                                 None if g is None else _ip(g), self.horizon, p, q, precision, seed)
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
