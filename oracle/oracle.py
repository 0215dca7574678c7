"""ctypes front end of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, ``__graft_entry__.smoke()`` and
the ``cpu_baseline`` leg of bench.py, always as the checker or the timed CPU
baseline.  Nothing in ``cugmres_amd`` imports it.

Arrays cross the boundary as interleaved (re, im) pairs, row-major, index
``i*N + j`` with i the row (the reference's C++ orientation,
B/CUDA_GMRES/GMRES_kernels.cuh:107-109).
"""
import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

_dp = ctypes.POINTER(ctypes.c_double)
_fp = ctypes.POINTER(ctypes.c_float)


class GmresOpts(ctypes.Structure):
    _fields_ = [("restart", ctypes.c_int), ("max_cycles", ctypes.c_int),
                ("tol", ctypes.c_double), ("adjoint", ctypes.c_int),
                ("reference_form", ctypes.c_int)]


class GmresInfo(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("cycles", ctypes.c_int),
                ("converged", ctypes.c_int), ("rel_residual", ctypes.c_double),
                ("beta0", ctypes.c_double)]


class _Op(ctypes.Structure):
    _fields_ = [("N", ctypes.c_int), ("M", ctypes.c_int), ("ghat", _dp),
                ("ghat32", _fp), ("tw", _dp)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _LIB
    if _LIB is None:
        path = os.path.join(_HERE, "liboracle.so")
        if not os.path.exists(path):
            build()
        L = ctypes.CDLL(path)
        L.orc_bessel01.argtypes = [ctypes.c_double] + [_dp] * 4
        L.orc_ref_wave_number.restype = ctypes.c_float
        L.orc_ref_wave_number.argtypes = [ctypes.c_int]
        L.orc_ref_green_block_f32.argtypes = [ctypes.c_int, _fp]
        L.orc_ref_gprep_f32.argtypes = [ctypes.c_int, _fp]
        L.orc_baseline_green_block.argtypes = [ctypes.c_int, ctypes.c_double, _dp]
        L.orc_shifted_green_block.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, _dp]
        L.orc_gprep_to_block.argtypes = [ctypes.c_int, _fp, _dp]
        L.orc_embed_block.argtypes = [ctypes.c_int, ctypes.c_int, _dp, _dp]
        L.orc_ref_cylinder_mask.argtypes = [ctypes.c_int, ctypes.POINTER(ctypes.c_uint8)]
        L.orc_ref_plane_wave_f32.argtypes = [ctypes.c_int, _fp]
        L.orc_plane_wave.argtypes = [ctypes.c_int, ctypes.c_double, ctypes.c_double, _dp]
        L.orc_fft_c128.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        L.orc_fft2_c128.argtypes = [ctypes.c_int, _dp, ctypes.c_int]
        L.orc_fft_c64.argtypes = [ctypes.c_int, _fp, ctypes.c_int]
        L.orc_fft2_c64.argtypes = [ctypes.c_int, _fp, ctypes.c_int]
        L.orc_op_init.argtypes = [ctypes.POINTER(_Op), ctypes.c_int, ctypes.c_int, _dp]
        L.orc_op_free.argtypes = [ctypes.POINTER(_Op)]
        L.orc_matvec_c128.argtypes = [ctypes.POINTER(_Op), _dp, _dp, _dp, ctypes.c_int]
        L.orc_matvec_c64.argtypes = [ctypes.POINTER(_Op), _fp, _fp, _fp, ctypes.c_int]
        L.orc_matvec_dense.argtypes = [ctypes.c_int, _dp, _dp, _dp, _dp, ctypes.c_int]
        L.orc_gmres_c128.argtypes = [ctypes.POINTER(_Op), _dp, _dp, _dp,
                                     ctypes.POINTER(GmresOpts), ctypes.POINTER(GmresInfo), _dp]
        L.orc_gmres_c64.argtypes = [ctypes.POINTER(_Op), _fp, _fp, _fp,
                                    ctypes.POINTER(GmresOpts), ctypes.POINTER(GmresInfo), _dp]
        L.orc_gmres_reference.argtypes = [ctypes.POINTER(_Op), _dp, _dp, _dp, ctypes.c_int,
                                          ctypes.c_int, _dp, _dp, _dp, _dp]
        L.orc_gmres_refine.argtypes = [ctypes.POINTER(_Op), _dp, _dp, _dp, ctypes.c_int,
                                       ctypes.c_double, ctypes.c_int, ctypes.c_double,
                                       ctypes.c_int, ctypes.POINTER(GmresInfo)]
        L.orc_now.restype = ctypes.c_double
        _LIB = L
    return _LIB


def _d(a):
    return a.ctypes.data_as(_dp)


def _f(a):
    return a.ctypes.data_as(_fp)


def bessel01(x):
    """(J0, J1, Y0, Y1) at x > 0, FP64."""
    out = [ctypes.c_double() for _ in range(4)]
    lib().orc_bessel01(float(x), *[ctypes.byref(v) for v in out])
    return tuple(v.value for v in out)


def ref_wave_number(N):
    return lib().orc_ref_wave_number(N)


def ref_green_block(N):
    a = np.zeros((N, N), np.complex64)
    lib().orc_ref_green_block_f32(N, _f(a))
    return a


def ref_gprep(N):
    a = np.zeros((2 * N - 1, 2 * N - 1), np.complex64)
    lib().orc_ref_gprep_f32(N, _f(a))
    return a


def baseline_green_block(N, kh):
    a = np.zeros((N, N), np.complex128)
    lib().orc_baseline_green_block(N, float(kh), _d(a))
    return a


def shifted_green_block(N, k, row_shift=0.5, col_shift=0.0):
    a = np.zeros((N, N), np.complex128)
    lib().orc_shifted_green_block(N, float(k), float(row_shift), float(col_shift), _d(a))
    return a


def gprep_to_block(N, gprep):
    g = np.ascontiguousarray(gprep, np.complex64).reshape(2 * N - 1, 2 * N - 1)
    blk = np.zeros((N, N), np.complex128)
    rc = lib().orc_gprep_to_block(N, _f(g), _d(blk))
    return blk, rc == 0


def embed_block(N, M, blk):
    b = np.ascontiguousarray(blk, np.complex128)
    E = np.zeros((M, M), np.complex128)
    lib().orc_embed_block(N, M, _d(b), _d(E))
    return E


def ref_cylinder_mask(N):
    m = np.zeros((N, N), np.uint8)
    lib().orc_ref_cylinder_mask(N, m.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)))
    return m


def ref_plane_wave(N):
    b = np.zeros((N, N), np.complex64)
    lib().orc_ref_plane_wave_f32(N, _f(b))
    return b


def plane_wave(N, kh, theta):
    b = np.zeros((N, N), np.complex128)
    lib().orc_plane_wave(N, float(kh), float(theta), _d(b))
    return b


def fft(x, sign=-1):
    a = np.array(x, dtype=np.complex128, copy=True)
    lib().orc_fft_c128(a.size, _d(a), sign)
    return a


def fft2(x, sign=-1):
    a = np.array(x, dtype=np.complex128, copy=True)
    assert a.ndim == 2 and a.shape[0] == a.shape[1]
    lib().orc_fft2_c128(a.shape[0], _d(a), sign)
    return a


def ref_chi(N):
    """Reference CHI = (EPSILON-1)*WAVE_NUMBER*WAVE_NUMBER with the macro's
    left-to-right float evaluation (B/CUDA_GMRES/GMRES_kernels.cuh:12,18-19)."""
    f = np.float32
    lam = f(N) / f(6.0)
    v = f(f(f(2.25) - f(1)) * f(2)) * f(3.14)
    v = f(v / lam)
    v = f(v * f(2))
    v = f(v * f(3.14))
    v = f(v / lam)
    return float(v)


def dense_matvec(N, blk, chi, x, mode=0):
    b = np.ascontiguousarray(blk, np.complex128)
    c = np.ascontiguousarray(chi, np.float64).reshape(-1)
    xx = np.ascontiguousarray(x, np.complex128).reshape(-1)
    y = np.zeros(N * N, np.complex128)
    lib().orc_matvec_dense(N, _d(b), _d(c), _d(xx), _d(y), mode)
    return y


class Op:
    """Circulant-embedded operator A = I - G diag(chi) on an N x N grid."""

    def __init__(self, N, blk, M=None):
        self.N = N
        self.M = M if M is not None else 2 * N
        self._blk = np.ascontiguousarray(blk, np.complex128)
        self._op = _Op()
        rc = lib().orc_op_init(ctypes.byref(self._op), N, self.M, _d(self._blk))
        if rc:
            raise ValueError("embedding size M must be >= 2N-1")

    def __del__(self):
        try:
            lib().orc_op_free(ctypes.byref(self._op))
        except Exception:
            pass

    @property
    def ghat(self):
        M = self.M
        return np.ctypeslib.as_array(self._op.ghat, shape=(M * M * 2,)).view(np.complex128).reshape(M, M).copy()

    def matvec(self, chi, x, mode=0):
        c = np.ascontiguousarray(chi, np.float64).reshape(-1)
        xx = np.ascontiguousarray(x, np.complex128).reshape(-1)
        y = np.zeros_like(xx)
        lib().orc_matvec_c128(ctypes.byref(self._op), _d(c), _d(xx), _d(y), mode)
        return y

    def matvec_c64(self, chi, x, mode=0):
        c = np.ascontiguousarray(chi, np.float32).reshape(-1)
        xx = np.ascontiguousarray(x, np.complex64).reshape(-1)
        y = np.zeros_like(xx)
        lib().orc_matvec_c64(ctypes.byref(self._op), _f(c), _f(xx), _f(y), mode)
        return y

    def gmres(self, chi, b, x0=None, restart=30, tol=1e-6, max_cycles=100, adjoint=False,
              history=False, precision="c128"):
        n = self.N * self.N
        if precision == "c128":
            c = np.ascontiguousarray(chi, np.float64).reshape(-1)
            bb = np.ascontiguousarray(b, np.complex128).reshape(-1)
            x = np.zeros(n, np.complex128) if x0 is None else np.array(x0, np.complex128).reshape(-1)
            conv, fnc = _d, lib().orc_gmres_c128
        else:
            c = np.ascontiguousarray(chi, np.float32).reshape(-1)
            bb = np.ascontiguousarray(b, np.complex64).reshape(-1)
            x = np.zeros(n, np.complex64) if x0 is None else np.array(x0, np.complex64).reshape(-1)
            conv, fnc = _f, lib().orc_gmres_c64
        o = GmresOpts(restart, max_cycles, tol, int(adjoint), 0)
        info = GmresInfo()
        hist = np.zeros(restart * (max_cycles + 1) + 2, np.float64)
        fnc(ctypes.byref(self._op), conv(c), conv(bb), conv(x), ctypes.byref(o), ctypes.byref(info),
            _d(hist) if history else None)
        res = dict(iterations=info.iterations, cycles=info.cycles, converged=bool(info.converged),
                   rel_residual=info.rel_residual, beta0=info.beta0)
        if history:
            res["history"] = hist[: info.iterations + 1].copy()
        return x, res

    def gmres_reference(self, chi, b, x0, maxiter, adjoint=False):
        n = self.N * self.N
        c = np.ascontiguousarray(chi, np.float64).reshape(-1)
        bb = np.ascontiguousarray(b, np.complex128).reshape(-1)
        x = np.array(x0, np.complex128).reshape(-1)
        H = np.zeros((maxiter + 1, maxiter), np.complex128)
        J = np.zeros((maxiter + 1, maxiter + 1), np.complex128)
        V = np.zeros((maxiter + 1, n), np.complex128)
        res = np.zeros(maxiter + 1, np.float64)
        lib().orc_gmres_reference(ctypes.byref(self._op), _d(c), _d(bb), _d(x), maxiter, int(adjoint),
                                  _d(H), _d(J), _d(V), _d(res))
        return x, dict(H=H, J=J, V=V, res=res)

    def gmres_refine(self, chi, b, x0=None, inner_restart=50, inner_tol=1e-5, inner_max_cycles=40,
                     outer_tol=1e-12, max_outer=30):
        n = self.N * self.N
        c = np.ascontiguousarray(chi, np.float64).reshape(-1)
        bb = np.ascontiguousarray(b, np.complex128).reshape(-1)
        x = np.zeros(n, np.complex128) if x0 is None else np.array(x0, np.complex128).reshape(-1)
        info = GmresInfo()
        lib().orc_gmres_refine(ctypes.byref(self._op), _d(c), _d(bb), _d(x), inner_restart, inner_tol,
                               inner_max_cycles, outer_tol, max_outer, ctypes.byref(info))
        return x, dict(iterations=info.iterations, outer=info.cycles, converged=bool(info.converged),
                       rel_residual=info.rel_residual)


def now():
    return lib().orc_now()
