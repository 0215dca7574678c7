"""ctypes binding of cugmres_amd/libcugmres.so (the C ABI in include/cugmres.h).

The library is built in-tree by ``python __graft_entry__.py`` / ``make -C
cugmres_amd``.  There is no fallback: if the shared object is missing every
entry point raises, so a GPU run can never silently use another path.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libcugmres.so")

_lib = None

c_float_p = ctypes.POINTER(ctypes.c_float)
c_double_p = ctypes.POINTER(ctypes.c_double)
c_void_p = ctypes.c_void_p

SUCCESS = 0
ERROR_INVALID_VALUE = 1
ERROR_DEVICE = 2
ERROR_ALLOC = 3
ERROR_NOT_CONVERGED = 4
ERROR_NONFINITE = 5
ERROR_UNSUPPORTED = 7
ERROR_NOT_SYMMETRIC = 8

OP_FORWARD = 0
OP_ADJOINT = 1
OP_GREEN = 2


class GmresParams(ctypes.Structure):
    _fields_ = [("restart", ctypes.c_int), ("max_cycles", ctypes.c_int),
                ("tol", ctypes.c_double), ("adjoint", ctypes.c_int), ("x0_zero", ctypes.c_int)]


class GmresResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("cycles", ctypes.c_int),
                ("converged", ctypes.c_int), ("rel_residual", ctypes.c_double),
                ("beta0", ctypes.c_double), ("bnorm", ctypes.c_double)]


class RefineParams(ctypes.Structure):
    _fields_ = [("inner_restart", ctypes.c_int), ("inner_max_cycles", ctypes.c_int),
                ("inner_tol", ctypes.c_double), ("outer_tol", ctypes.c_double),
                ("max_outer", ctypes.c_int)]


class RefineResult(ctypes.Structure):
    _fields_ = [("outer", ctypes.c_int), ("inner_iterations", ctypes.c_int),
                ("converged", ctypes.c_int), ("rel_residual", ctypes.c_double)]


class FgmresParams(ctypes.Structure):
    _fields_ = [("restart", ctypes.c_int), ("max_cycles", ctypes.c_int), ("tol", ctypes.c_double),
                ("x0_zero", ctypes.c_int), ("stall_ratio", ctypes.c_double)]


class FgmresResult(ctypes.Structure):
    _fields_ = [("iterations", ctypes.c_int), ("cycles", ctypes.c_int), ("converged", ctypes.c_int),
                ("rel_residual", ctypes.c_double), ("beta0", ctypes.c_double), ("bnorm", ctypes.c_double),
                ("precond_cycles", ctypes.c_int), ("fell_back", ctypes.c_int)]


class Precond(ctypes.Structure):
    """cgm_precond: the tile inverses and/or the network (NULL = absent)."""
    _fields_ = [("tiles", c_void_p), ("net", c_void_p)]


# name -> (restype, argtypes); every symbol declared in include/cugmres.h
SIGNATURES = {
    "cgm_status_string": (ctypes.c_char_p, [ctypes.c_int]),
    "cgm_last_error": (ctypes.c_char_p, []),
    "cgm_version": (None, [ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int)]),
    "cgm_create": (ctypes.c_int, [ctypes.POINTER(c_void_p), ctypes.c_int, ctypes.c_int]),
    "cgm_destroy": (ctypes.c_int, [c_void_p]),
    "cgm_set_stream": (ctypes.c_int, [c_void_p, c_void_p]),
    "cgm_get_embedding": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "cgm_set_arnoldi_path": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "cgm_autotune": (ctypes.c_int, [c_void_p]),
    "cgm_get_tiles": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_int), c_float_p]),
    "cgm_get_radix_plan": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_int), c_float_p, c_float_p]),
    "cgm_set_radix_plan": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_int)]),
    "cgm_spectrum_size": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_size_t)]),
    "cgm_spectrum_from_gprep": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "cgm_spectrum_from_block": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p]),
    "cgm_spectrum_reference": (ctypes.c_int, [c_void_p, ctypes.c_double, ctypes.c_double,
                                              ctypes.c_double, c_void_p, c_void_p]),
    "cgm_spectrum_helmholtz": (ctypes.c_int, [c_void_p, ctypes.c_double, c_void_p, c_void_p]),
    "cgm_matvec": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int]),
    "cgm_matvec_batched": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_long, c_void_p, c_void_p,
                                          ctypes.c_int, ctypes.c_int]),
    "cgm_matvec_batched_noalloc": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_long, c_void_p,
                                                  c_void_p, ctypes.c_int, ctypes.c_int]),
    "cgm_noalloc_scratch_bytes": (ctypes.c_size_t, []),
    "cgm_matvec_c128": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int]),
    "cgm_gmres": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                 ctypes.POINTER(GmresParams), ctypes.POINTER(GmresResult), c_void_p]),
    "cgm_set_batched_graphs": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "cgm_gmres_batched": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_long, c_void_p, c_void_p,
                                         ctypes.c_int, ctypes.POINTER(GmresParams), ctypes.POINTER(GmresResult),
                                         c_void_p]),
    "cgm_gmres_refine": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                        ctypes.POINTER(RefineParams), ctypes.POINTER(RefineResult)]),
    "cgm_gmres_refine_pc": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                           ctypes.POINTER(Precond), ctypes.POINTER(RefineParams),
                                           ctypes.POINTER(RefineResult)]),
    "cgm_fno_num_params": (ctypes.c_int, [ctypes.POINTER(ctypes.c_size_t)]),
    "cgm_fno_create": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.c_float, ctypes.POINTER(c_void_p)]),
    "cgm_fno_destroy": (ctypes.c_int, [c_void_p]),
    "cgm_fno_set_params": (ctypes.c_int, [c_void_p, c_void_p]),
    "cgm_fno_forward": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_int]),
    "cgm_fno_backward": (ctypes.c_int, [c_void_p, c_void_p, c_void_p]),
    "cgm_fno_profile": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_int,
                                       c_float_p]),
    "cgm_fgmres": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, c_void_p,
                                  ctypes.POINTER(FgmresParams), ctypes.POINTER(FgmresResult), c_void_p]),
    "cgm_fgmres_batched": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_long, c_void_p, c_void_p,
                                          c_void_p, ctypes.c_int, ctypes.POINTER(FgmresParams),
                                          ctypes.POINTER(FgmresResult), c_void_p]),
    "cgm_block_helmholtz": (ctypes.c_int, [c_void_p, ctypes.c_double, c_void_p]),
    "cgm_block_reference": (ctypes.c_int, [c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                           c_void_p]),
    "cgm_tiles_create": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.c_int, ctypes.POINTER(c_void_p)]),
    "cgm_tiles_destroy": (ctypes.c_int, [c_void_p]),
    "cgm_tiles_memory": (ctypes.c_int, [c_void_p, ctypes.POINTER(ctypes.c_size_t)]),
    "cgm_tiles_apply_bytes": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]),
    "cgm_tiles_factor": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_long, ctypes.c_int]),
    "cgm_tiles_apply": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_int]),
    "cgm_set_flexible_fp64": (ctypes.c_int, [c_void_p, c_void_p]),
    "cgm_fgmres_pc": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.POINTER(Precond),
                                     ctypes.POINTER(FgmresParams), ctypes.POINTER(FgmresResult), c_void_p]),
    "cgm_fgmres_pc_batched": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.c_long, c_void_p, c_void_p,
                                             ctypes.POINTER(Precond), ctypes.c_int, ctypes.POINTER(FgmresParams),
                                             ctypes.POINTER(FgmresResult), c_void_p]),
    "cgm_reference_plane_wave": (ctypes.c_int, [c_void_p, c_void_p]),
    "cgm_reference_chi": (ctypes.c_float, [ctypes.c_int]),
    "cgm_set_mask_contrast": (ctypes.c_int, [c_void_p, ctypes.c_float]),
    "cgm_n_timestamps": (ctypes.c_uint, [ctypes.c_uint]),
    "cgm_Fast_GMRES_with_CUDA": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, ctypes.POINTER(c_void_p),
                                                ctypes.POINTER(ctypes.c_uint), c_void_p, ctypes.c_float,
                                                ctypes.c_bool, ctypes.c_uint, ctypes.POINTER(ctypes.c_bool),
                                                ctypes.c_uint, c_void_p, ctypes.c_uint]),
    "cgm_G_x_fft_matvec": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_uint]),
    "cgm_G_x_fft_matvec_T": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_uint]),
    "cgm_cuSolve_LES": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_int, c_void_p, c_void_p, c_void_p,
                                       c_void_p]),
    "cgm_malloc": (ctypes.c_int, [ctypes.POINTER(c_void_p), ctypes.c_size_t]),
    "cgm_free": (ctypes.c_int, [c_void_p]),
    "cgm_memcpy_h2d": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_size_t]),
    "cgm_memcpy_d2h": (ctypes.c_int, [c_void_p, c_void_p, ctypes.c_size_t]),
    "cgm_synchronize": (ctypes.c_int, [c_void_p]),
    "cgm_bessel01": (None, [ctypes.c_double] + [c_double_p] * 4),
    "cgm_green_values": (ctypes.c_int, [c_void_p, ctypes.c_size_t, ctypes.c_double, c_void_p]),
    "cgm_longrec_timing": (ctypes.c_int, [c_void_p, ctypes.c_int]),
    "cgm_longrec_timing_read": (ctypes.c_int, [c_void_p, c_void_p]),
    "cgm_profile_tiles": (ctypes.c_int, [c_void_p, ctypes.c_int, c_float_p]),
    "cgm_profile_longrec": (ctypes.c_int, [c_void_p, ctypes.c_int, ctypes.c_int, c_float_p]),
    "cgm_profile_step": (ctypes.c_int, [c_void_p, c_void_p, c_void_p, c_void_p, ctypes.c_int, ctypes.c_int,
                                        c_float_p]),
}


class CgmError(RuntimeError):
    def __init__(self, status, where):
        self.status = status
        msg = lib().cgm_last_error().decode(errors="replace")
        name = lib().cgm_status_string(status).decode()
        super().__init__(f"{where}: {name} ({status}) {msg}")


def lib():
    """Load the shared library (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(f"{LIB_PATH} is missing: build it with `make -C cugmres_amd` "
                              "or `python __graft_entry__.py`")
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in SIGNATURES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(status, where, ok=(SUCCESS,)):
    if status not in ok:
        raise CgmError(status, where)
    return status
