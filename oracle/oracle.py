"""ctypes front end of the FP64 CPU oracle (liboracle.so) and of the
reference's own dependency-free build (_ref/libssa_ref.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py, always as the checker or the timed CPU baseline.
The product library never loads it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_dp = ctypes.POINTER(ctypes.c_double)
_ip = ctypes.POINTER(ctypes.c_int)


def _d(a):
    return a.ctypes.data_as(_dp)


def _i(a):
    return a.ctypes.data_as(_ip)


def load():
    lib = ctypes.CDLL(os.path.join(_HERE, "liboracle.so"))
    vp = ctypes.c_void_p
    lib.orc_new.restype = vp
    lib.orc_delete.argtypes = [vp]
    lib.orc_init.argtypes = [vp, _dp, ctypes.c_int, ctypes.c_int]
    lib.orc_update_signal.argtypes = [vp, _dp]
    lib.orc_matvec.argtypes = [vp, _dp, _dp]
    lib.orc_matvec_T.argtypes = [vp, _dp, _dp]
    lib.orc_matvec_block.argtypes = [vp, _dp, _dp, ctypes.c_int, ctypes.c_int]
    lib.orc_decompose.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_reconstruct.argtypes = [vp, _ip, ctypes.c_int, _dp]
    lib.orc_wcorr.argtypes = [vp, _dp]
    lib.orc_lrf.argtypes = [vp, _ip, ctypes.c_int, _dp, _dp]
    lib.orc_forecast.argtypes = [vp, _ip, ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_forecast_full.argtypes = [vp, _ip, ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_vforecast.argtypes = [vp, _ip, ctypes.c_int, _dp, ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_cadzow.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp, _dp]
    lib.orc_cadzow_inplace.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_uint64, _dp]
    lib.orc_gapfill.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_int, ctypes.c_double, _dp]
    lib.orc_gapfill_simple.argtypes = [_dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_variance_explained.argtypes = [vp, ctypes.c_int, ctypes.c_int]
    lib.orc_variance_explained.restype = ctypes.c_double
    lib.orc_dims.argtypes = [vp, _ip]
    lib.orc_get.argtypes = [vp, _dp, _dp, _dp]
    lib.orc_fft.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
    lib.orc_philox4x32_10.argtypes = [ctypes.POINTER(ctypes.c_uint32), ctypes.POINTER(ctypes.c_uint32),
                                      ctypes.POINTER(ctypes.c_uint32)]
    lib.orc_philox_gaussian.argtypes = [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_qr_q.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_jacobi_svd.argtypes = [_dp, ctypes.c_int, ctypes.c_int, _dp, _dp]
    lib.orc_mssa_new.restype = vp
    lib.orc_mssa_delete.argtypes = [vp]
    lib.orc_mssa_init.argtypes = [vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int]
    lib.orc_mssa_matvec.argtypes = [vp, _dp, _dp]
    lib.orc_mssa_matvec_T.argtypes = [vp, _dp, _dp]
    lib.orc_mssa_decompose.argtypes = [vp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp]
    lib.orc_mssa_reconstruct.argtypes = [vp, ctypes.c_int, _ip, ctypes.c_int, _dp]
    lib.orc_mssa_get.argtypes = [vp, _dp, _dp, _dp]
    lib.orc_mssa_contributions.argtypes = [vp, _dp]
    return lib


_LIB = None


def lib():
    global _LIB
    if _LIB is None:
        _LIB = load()
    return _LIB


def philox_gaussian(seed, call, rows, cols):
    out = np.zeros(rows * cols, dtype=np.float64)
    lib().orc_philox_gaussian(seed, call, rows, cols, _d(out))
    return out.reshape(cols, rows).T.copy()   # rows x cols


def philox4x32_10(ctr, key):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def fft(z, sign=-1):
    z = np.ascontiguousarray(z, dtype=np.complex128).copy()
    buf = z.view(np.float64)
    lib().orc_fft(_d(buf), len(z), sign)
    return buf.view(np.complex128).copy()


class OracleSSA:
    """FP64 CPU SSA following MKL/ssa_opt.h; Omega is an explicit input."""

    def __init__(self, x, L):
        self._l = lib()
        self._p = self._l.orc_new()
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.N = len(self.x)
        if self._l.orc_init(self._p, _d(self.x), self.N, L) != 0:
            raise ValueError("oracle init failed")
        d = np.zeros(4, dtype=np.int32)
        self._l.orc_dims(self._p, _i(d))
        self.N, self.L, self.K, self.n = (int(v) for v in d)
        self.k = 0

    def __del__(self):
        if getattr(self, "_p", None):
            self._l.orc_delete(self._p)
            self._p = None

    def update_signal(self, x):
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        assert self._l.orc_update_signal(self._p, _d(self.x)) == 0

    def matvec(self, v):
        y = np.zeros(self.L)
        self._l.orc_matvec(self._p, _d(np.ascontiguousarray(v, dtype=np.float64)), _d(y))
        return y

    def matvec_T(self, u):
        z = np.zeros(self.K)
        self._l.orc_matvec_T(self._p, _d(np.ascontiguousarray(u, dtype=np.float64)), _d(z))
        return z

    def matvec_block(self, V, transpose=False):
        V = np.asfortranarray(V, dtype=np.float64)
        b = V.shape[1]
        out = np.zeros((self.K if transpose else self.L, b), order="F")
        self._l.orc_matvec_block(self._p, _d(V), _d(out), b, int(transpose))
        return out

    def decompose(self, k, p, q, omega):
        omega = np.asfortranarray(omega, dtype=np.float64)
        rc = self._l.orc_decompose(self._p, k, p, q, _d(omega))
        if rc != 0:
            raise RuntimeError("oracle decompose failed")
        d = np.zeros(4, dtype=np.int32)
        self.k = self._l.orc_dims(self._p, _i(d))
        return self

    def factors(self):
        U = np.zeros((self.L, self.k), order="F")
        V = np.zeros((self.K, self.k), order="F")
        s = np.zeros(self.k)
        self._l.orc_get(self._p, _d(U), _d(V), _d(s))
        return U, V, s

    @property
    def sigma(self):
        return self.factors()[2]

    def reconstruct(self, group):
        g = np.ascontiguousarray(group, dtype=np.int32)
        out = np.zeros(self.N)
        if self._l.orc_reconstruct(self._p, _i(g), len(g), _d(out)) != 0:
            raise RuntimeError("oracle reconstruct failed")
        return out

    def wcorr(self):
        W = np.zeros((self.k, self.k))
        assert self._l.orc_wcorr(self._p, _d(W)) == 0
        return W

    def lrf(self, group):
        g = np.ascontiguousarray(group, dtype=np.int32)
        R = np.zeros(self.L - 1)
        nu = np.zeros(1)
        rc = self._l.orc_lrf(self._p, _i(g), len(g), _d(R), _d(nu))
        return rc, R, float(nu[0])

    def forecast(self, group, h):
        g = np.ascontiguousarray(group, dtype=np.int32)
        out = np.zeros(h)
        if self._l.orc_forecast(self._p, _i(g), len(g), h, _d(out)) != 0:
            raise RuntimeError("oracle forecast failed")
        return out

    def forecast_full(self, group, h):
        g = np.ascontiguousarray(group, dtype=np.int32)
        out = np.zeros(self.N + h)
        if self._l.orc_forecast_full(self._p, _i(g), len(g), h, _d(out)) != 0:
            raise RuntimeError("oracle forecast_full failed")
        return out

    def vforecast(self, group, h, base=None):
        """V-forecast from the group's reconstruction, or from the last L-1
        values of `base` (MKL/ssa_opt.h:2906-2965, :3016-3078)."""
        g = np.ascontiguousarray(group, dtype=np.int32)
        out = np.zeros(h)
        if base is None:
            rc = self._l.orc_vforecast(self._p, _i(g), len(g), None, 0, h, _d(out))
        else:
            b = np.ascontiguousarray(base, dtype=np.float64)
            rc = self._l.orc_vforecast(self._p, _i(g), len(g), _d(b), len(b), h, _d(out))
        if rc != 0:
            raise RuntimeError("oracle vforecast failed")
        return out

    def cadzow_inplace(self, rank, max_iter, tol, counter0=0):
        """Cadzow on this context's signal (MKL/ssa_opt.h:3319-3440); the
        decompositions use Philox counters counter0, counter0 + 1, ..."""
        r = np.zeros(3)
        it = self._l.orc_cadzow_inplace(self._p, rank, max_iter, tol, counter0, _d(r))
        if it < 0:
            raise RuntimeError("oracle cadzow_inplace failed")
        return it, float(r[1]), bool(r[2])

    def variance_explained(self, start, end):
        return self._l.orc_variance_explained(self._p, start, end)


def cadzow(x, L, rank, max_iter=20, tol=1e-9):
    """Cadzow iterations (MKL/ssa_opt.h:3165-3292): (signal, iterations, final_diff, converged)."""
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(len(x))
    r = np.zeros(3)
    it = lib().orc_cadzow(_d(x), len(x), L, rank, max_iter, tol, _d(out), _d(r))
    if it < 0:
        raise RuntimeError("oracle cadzow failed")
    return out, it, float(r[1]), bool(r[2])


def gapfill(x, L, rank, max_iter=50, tol=1e-6):
    """Iterative gap filling (MKL/ssa_opt.h:3498-3683) of the NaNs of a copy of x:
    (filled, iterations, final_diff, converged, n_gaps)."""
    y = np.array(x, dtype=np.float64)
    r = np.zeros(4)
    if lib().orc_gapfill(_d(y), len(y), L, rank, max_iter, tol, _d(r)) != 0:
        raise RuntimeError("oracle gapfill failed")
    return y, int(r[0]), float(r[1]), bool(r[2]), int(r[3])


def gapfill_simple(x, L, rank):
    """Forecast/backcast gap filling (MKL/ssa_opt.h:3685-3877) of a copy of x."""
    y = np.array(x, dtype=np.float64)
    r = np.zeros(4)
    if lib().orc_gapfill_simple(_d(y), len(y), L, rank, _d(r)) != 0:
        raise RuntimeError("oracle gapfill_simple failed")
    return y, int(r[3])


class OracleMSSA:
    def __init__(self, X, L):
        self._l = lib()
        self._p = self._l.orc_mssa_new()
        X = np.ascontiguousarray(X, dtype=np.float64)
        self.M, self.N = X.shape
        self.L = L
        self.K = self.N - L + 1
        if self._l.orc_mssa_init(self._p, _d(X), self.M, self.N, L) != 0:
            raise ValueError("oracle mssa init failed")
        self.k = 0

    def __del__(self):
        if getattr(self, "_p", None):
            self._l.orc_mssa_delete(self._p)
            self._p = None

    def matvec(self, v):
        y = np.zeros(self.M * self.L)
        self._l.orc_mssa_matvec(self._p, _d(np.ascontiguousarray(v, dtype=np.float64)), _d(y))
        return y

    def matvec_T(self, u):
        z = np.zeros(self.K)
        self._l.orc_mssa_matvec_T(self._p, _d(np.ascontiguousarray(u, dtype=np.float64)), _d(z))
        return z

    def decompose(self, k, p, q, omega):
        omega = np.asfortranarray(omega, dtype=np.float64)
        if self._l.orc_mssa_decompose(self._p, k, p, q, _d(omega)) != 0:
            raise RuntimeError("oracle mssa decompose failed")
        b = min(k + (p if p > 0 else 8), min(self.M * self.L, self.K))
        self.k = min(k, b)
        return self

    def factors(self):
        ML = self.M * self.L
        U = np.zeros((ML, self.k), order="F")
        V = np.zeros((self.K, self.k), order="F")
        s = np.zeros(self.k)
        self._l.orc_mssa_get(self._p, _d(U), _d(V), _d(s))
        return U, V, s

    def reconstruct(self, series, group):
        g = np.ascontiguousarray(group, dtype=np.int32)
        out = np.zeros(self.N)
        if self._l.orc_mssa_reconstruct(self._p, series, _i(g), len(g), _d(out)) != 0:
            raise RuntimeError("oracle mssa reconstruct failed")
        return out

    def contributions(self):
        c = np.zeros(self.M * self.k)
        self._l.orc_mssa_contributions(self._p, _d(c))
        return c.reshape(self.M, self.k)


# --------------------------------------------------------------------------
# the reference's dependency-free build (MKL/ssa_opt_ref.h), q = 0 only
# --------------------------------------------------------------------------

class _SSARefStruct(ctypes.Structure):
    # field order of SSA_Opt_Ref, MKL/ssa_opt_ref.h:90-120
    _fields_ = [("N", ctypes.c_int), ("L", ctypes.c_int), ("K", ctypes.c_int), ("fft_len", ctypes.c_int),
                ("fft_x", _dp), ("inv_diag_count", _dp),
                ("ws_fft1", _dp), ("ws_fft2", _dp), ("ws_fft3", _dp),
                ("ws_u", _dp), ("ws_v", _dp), ("ws_proj", _dp),
                ("U", _dp), ("V", _dp), ("sigma", _dp), ("eigenvalues", _dp),
                ("n_components", ctypes.c_int), ("initialized", ctypes.c_bool), ("decomposed", ctypes.c_bool),
                ("total_variance", ctypes.c_double), ("rng_state", ctypes.c_uint64),
                ("_pad", ctypes.c_char * 256)]


REF_SO = os.path.join(_HERE, "_ref", "libssa_ref.so")


def ref_available():
    return os.path.exists(REF_SO)


def ref_xorshift_omega(K, b, state=0x123456789ABCDEF0):
    """The Gaussian matrix ssa_opt_ref_decompose_randomized draws from a fresh
    context: Marsaglia xorshift64 (13, 7, 17) uniforms on [-0.5, 0.5) and
    Box-Muller with the cosine branch, column-major K x b
    (MKL/ssa_opt_ref.h:305-323, :1114-1115, seed :817)."""
    m = (1 << 64) - 1
    x = state
    out = np.empty(K * b)

    def rnd():
        nonlocal x
        x ^= (x << 13) & m
        x ^= x >> 7
        x ^= (x << 17) & m
        return (x >> 11) * (1.0 / 9007199254740992.0) - 0.5

    for i in range(K * b):
        u1 = rnd() + 0.5
        u2 = rnd() + 0.5
        if u1 < 1e-10:
            u1 = 1e-10
        out[i] = np.sqrt(-2.0 * np.log(u1)) * np.cos(2.0 * 3.14159265358979323846 * u2)
    return out.reshape(b, K).T.copy()


class RefSSA:
    """The reference's own ssa_opt_ref.h, compiled from source into _ref/."""

    def __init__(self, x, L):
        self._l = ctypes.CDLL(REF_SO)
        self._l.ssa_opt_ref_init.argtypes = [ctypes.c_void_p, _dp, ctypes.c_int, ctypes.c_int]
        self._l.ssa_opt_ref_decompose_randomized.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_int]
        self._l.ssa_opt_ref_reconstruct.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int, _dp]
        self._l.ssa_opt_ref_wcorr_matrix.argtypes = [ctypes.c_void_p, _dp]
        self._l.ssa_opt_ref_free.argtypes = [ctypes.c_void_p]
        self._l.ssa_opt_ref_compute_lrf.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int, _dp, _dp]
        self._l.ssa_opt_ref_forecast.argtypes = [ctypes.c_void_p, _ip, ctypes.c_int, ctypes.c_int, _dp]
        self.s = _SSARefStruct()
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.N = len(self.x)
        if self._l.ssa_opt_ref_init(ctypes.byref(self.s), _d(self.x), self.N, L) != 0:
            raise ValueError("ref init failed")

    def __del__(self):
        if getattr(self, "_l", None) is not None:
            self._l.ssa_opt_ref_free(ctypes.byref(self.s))

    def decompose_randomized(self, k, p):
        assert self._l.ssa_opt_ref_decompose_randomized(ctypes.byref(self.s), k, p) == 0
        self.k = self.s.n_components
        return self

    @property
    def sigma(self):
        return np.ctypeslib.as_array(self.s.sigma, shape=(self.k,)).copy()

    def factors(self):
        U = np.ctypeslib.as_array(self.s.U, shape=(self.k * self.s.L,)).reshape(self.k, self.s.L).T.copy()
        V = np.ctypeslib.as_array(self.s.V, shape=(self.k * self.s.K,)).reshape(self.k, self.s.K).T.copy()
        return U, V, self.sigma

    def reconstruct(self, group):
        g = np.ascontiguousarray(group, dtype=np.int32)
        out = np.zeros(self.N)
        assert self._l.ssa_opt_ref_reconstruct(ctypes.byref(self.s), _i(g), len(g), _d(out)) == 0
        return out

    def wcorr(self):
        W = np.zeros((self.k, self.k))
        assert self._l.ssa_opt_ref_wcorr_matrix(ctypes.byref(self.s), _d(W)) == 0
        return W

    def forecast(self, group, h):
        g = np.ascontiguousarray(group, dtype=np.int32)
        out = np.zeros(self.N + h)
        rc = self._l.ssa_opt_ref_forecast(ctypes.byref(self.s), _i(g), len(g), h, _d(out))
        return rc, out


MKL_SO = os.path.join(_HERE, "libmkl_baseline.so")
MKL_RT = os.environ.get("SSA_MKL_RT", "/opt/conda/lib/libmkl_rt.so.1")
_mkl = None


def mkl_lib():
    """libmkl_baseline.so with MKL's runtime loaded, or None where either is
    absent (then the baseline falls back to the oracle port)."""
    global _mkl
    if _mkl is not None:
        return _mkl or None
    _mkl = False
    if not (os.path.exists(MKL_SO) and os.path.exists(MKL_RT)):
        return None
    lib = ctypes.CDLL(MKL_SO)
    vp = ctypes.c_void_p
    lib.mklb_load.argtypes = [ctypes.c_char_p]
    lib.mklb_set_threads.argtypes = [ctypes.c_int]
    lib.mklb_version.argtypes = [ctypes.c_char_p, ctypes.c_int]
    lib.mklb_create.restype = vp
    lib.mklb_create.argtypes = [_dp, ctypes.c_int, ctypes.c_int]
    lib.mklb_free.argtypes = [vp]
    lib.mklb_decompose.argtypes = [vp, _dp, ctypes.c_int, ctypes.c_int, ctypes.c_int, _dp, _dp, _dp]
    lib.mklb_reconstruct.argtypes = [vp, _dp, _dp, _dp, _ip, ctypes.c_int, _dp]
    lib.mklb_wcorr.argtypes = [vp, _dp, _dp, ctypes.c_int, _dp]
    if lib.mklb_load(MKL_RT.encode()) != 0:
        return None
    _mkl = lib
    return lib


def mkl_version():
    lib = mkl_lib()
    if lib is None:
        return None
    buf = ctypes.create_string_buffer(256)
    lib.mklb_version(buf, 256)
    return buf.value.decode(errors="replace").strip()


class MklSSA:
    """One series through the reference's MKL call sequence (mkl_baseline.c):
    decompose(k, p, q, omega) with the caller's K x (k+p) test matrix, then
    reconstructions.  The CPU baseline, checked against the oracle."""

    def __init__(self, x, L):
        self._l = mkl_lib()
        if self._l is None:
            raise RuntimeError("MKL runtime not available")
        self.x = np.ascontiguousarray(x, dtype=np.float64)
        self.N, self.L = len(self.x), L
        self.K = self.N - L + 1
        self._h = self._l.mklb_create(_d(self.x), self.N, L)
        if not self._h:
            raise RuntimeError("MKL context failed")

    def __del__(self):
        if getattr(self, "_h", None):
            self._l.mklb_free(self._h)
            self._h = None

    def decompose(self, k, p, q, omega):
        b = omega.shape[1]
        om = np.asfortranarray(omega, dtype=np.float64)
        self.k = k
        self.sigma = np.zeros(k)
        self.U = np.zeros((self.L, k), order="F")
        self.V = np.zeros((self.K, k), order="F")
        if self._l.mklb_decompose(self._h, _d(om), k, b, q, _d(self.sigma), _d(self.U), _d(self.V)) != 0:
            raise RuntimeError("MKL decomposition failed")
        return self

    def reconstruct(self, group):
        g = np.ascontiguousarray(group, dtype=np.int32)
        out = np.zeros(self.N)
        if self._l.mklb_reconstruct(self._h, _d(self.sigma), _d(self.U), _d(self.V), _i(g), len(g), _d(out)) != 0:
            raise RuntimeError("MKL reconstruction failed")
        return out

    def wcorr(self):
        W = np.zeros((self.k, self.k))
        if self._l.mklb_wcorr(self._h, _d(self.U), _d(self.V), self.k, _d(W)) != 0:
            raise RuntimeError("MKL W-correlation failed")
        return W
