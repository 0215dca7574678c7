"""Device-level Python front end of the C ABI.

PyTorch-ROCm is used only as the device-memory and stream plumbing: vectors
are ``torch`` tensors on ``cuda:<device>`` whose ``data_ptr()`` is handed to
the C ABI, and the context is bound to torch's current stream.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

_c = ctypes


def _ptr(t):
    return _c.c_void_p(t.data_ptr()) if t is not None else None


class Context:
    """One solver context: grid size N, embedding M, one device, one stream."""

    def __init__(self, N, device=0):
        import torch
        self.torch = torch
        self.N = int(N)
        self.device = torch.device("cuda", device)
        p = _c.c_void_p()
        check(lib().cgm_create(_c.byref(p), self.N, device), "cgm_create")
        self._ptr = p
        stream = torch.cuda.current_stream(self.device).cuda_stream
        check(lib().cgm_set_stream(self._ptr, _c.c_void_p(stream)), "cgm_set_stream")
        m = _c.c_int()
        check(lib().cgm_get_embedding(self._ptr, _c.byref(m)), "cgm_get_embedding")
        self.M = m.value
        s = _c.c_size_t()
        check(lib().cgm_spectrum_size(self._ptr, _c.byref(s)), "cgm_spectrum_size")
        self.spectrum_len = s.value

    def close(self):
        if getattr(self, "_ptr", None) is not None:
            lib().cgm_destroy(self._ptr)
            self._ptr = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._ptr

    def tiles(self):
        """(rows per work-group of pass 1, columns of pass 2, rows of pass 3) and
        their tuned times in us, as chosen by the autotuner."""
        t = (_c.c_int * 3)()
        us = (_c.c_float * 3)()
        check(lib().cgm_get_tiles(self._ptr, t, us), "cgm_get_tiles")
        return tuple(t), tuple(us)

    def autotune(self):
        check(lib().cgm_autotune(self._ptr), "cgm_autotune")

    def radix_plan(self):
        """The autotuner's radix plan at M = 1024: {'wide_from_batch': [pass 1, 2, 3],
        'wave_us' / 'lds_us': [pass][batch 1, 4, 16] for the one-wave register kernel
        and the LDS-staged radix-8 kernel}."""
        wb = (_c.c_int * 3)()
        w = (_c.c_float * 9)()
        l_ = (_c.c_float * 9)()
        check(lib().cgm_get_radix_plan(self._ptr, wb, w, l_), "cgm_get_radix_plan")
        return {"wide_from_batch": list(wb), "wave_us": [[round(w[3 * q + k], 2) for k in range(3)] for q in range(3)],
                "lds_us": [[round(l_[3 * q + k], 2) for k in range(3)] for q in range(3)]}

    def set_radix_plan(self, wide_from_batch):
        """wide_from_batch: one int for all passes or a list of three."""
        v = wide_from_batch if isinstance(wide_from_batch, (list, tuple)) else [wide_from_batch] * 3
        check(lib().cgm_set_radix_plan(self._ptr, (_c.c_int * 3)(*[int(x) for x in v])), "cgm_set_radix_plan")

    def set_flexible_fp64(self, gamma64):
        """Long-recurrence flexible solves form A z_j and their residuals with
        the FP64 passes and this complex128 spectrum (None: complex64 passes).
        The tensor is kept alive by the context object."""
        self._flex_g64 = gamma64
        check(lib().cgm_set_flexible_fp64(self._ptr, _ptr(gamma64)), "cgm_set_flexible_fp64")

    def longrec_timing(self, enable=True):
        """Live HIP-event timing of the long-recurrence projection and update
        kernels in the solves that follow (sums zeroed)."""
        check(lib().cgm_longrec_timing(self._ptr, int(bool(enable))), "cgm_longrec_timing")

    def longrec_timing_read(self):
        """{'proj'|'upd'|'tile': {'ms', 'bytes', 'launches'}} summed since longrec_timing()."""
        out = (_c.c_double * 9)()
        check(lib().cgm_longrec_timing_read(self._ptr, out), "cgm_longrec_timing_read")
        return {k: {"ms": out[3 * i], "bytes": out[3 * i + 1], "launches": int(out[3 * i + 2])}
                for i, k in enumerate(("proj", "upd", "tile"))}

    def set_arnoldi_path(self, path):
        """0: fused lagged-CGS2 steps (default), 1: separate two-pass CGS2 kernels."""
        check(lib().cgm_set_arnoldi_path(self._ptr, int(path)), "cgm_set_arnoldi_path")

    # ---- tensors --------------------------------------------------------
    def empty(self, dtype="c64"):
        t = self.torch
        dt = {"c64": t.complex64, "c128": t.complex128, "f32": t.float32, "f64": t.float64}[dtype]
        return t.empty(self.N * self.N, dtype=dt, device=self.device)

    def to_device(self, a, dtype="c64"):
        t = self.torch
        dt = {"c64": t.complex64, "c128": t.complex128, "f32": t.float32, "f64": t.float64}[dtype]
        return t.as_tensor(np.ascontiguousarray(a).reshape(-1)).to(device=self.device, dtype=dt).contiguous()

    def _spectra(self, want32, want64):
        t = self.torch
        g32 = t.empty(self.spectrum_len, dtype=t.complex64, device=self.device) if want32 else None
        g64 = t.empty(self.spectrum_len, dtype=t.complex128, device=self.device) if want64 else None
        return g32, g64

    # ---- spectra ----------------------------------------------------------
    def spectrum_helmholtz(self, kh, fp64=False):
        g32, g64 = self._spectra(True, fp64)
        check(lib().cgm_spectrum_helmholtz(self._ptr, float(kh), _ptr(g32), _ptr(g64)),
              "cgm_spectrum_helmholtz")
        return (g32, g64) if fp64 else g32

    def spectrum_reference(self, k, row_shift=0.5, col_shift=0.0, fp64=False):
        g32, g64 = self._spectra(True, fp64)
        check(lib().cgm_spectrum_reference(self._ptr, float(k), float(row_shift), float(col_shift),
                                           _ptr(g32), _ptr(g64)), "cgm_spectrum_reference")
        return (g32, g64) if fp64 else g32

    def spectrum_from_gprep(self, gprep):
        g = self.to_device(np.asarray(gprep, np.complex64), "c64")
        g32, _ = self._spectra(True, False)
        check(lib().cgm_spectrum_from_gprep(self._ptr, _ptr(g), _ptr(g32)), "cgm_spectrum_from_gprep")
        return g32

    def spectrum_from_block(self, block, fp64=False):
        b = self.to_device(np.asarray(block, np.complex128), "c128")
        g32, g64 = self._spectra(True, fp64)
        check(lib().cgm_spectrum_from_block(self._ptr, _ptr(b), _ptr(g32), _ptr(g64)),
              "cgm_spectrum_from_block")
        return (g32, g64) if fp64 else g32

    def block_helmholtz(self, kh):
        """The N x N generator block g[|di|, |dj|] of the benchmark operator
        (complex128, device)."""
        t = self.torch
        blk = t.empty(self.N * self.N, dtype=t.complex128, device=self.device)
        check(lib().cgm_block_helmholtz(self._ptr, float(kh), _ptr(blk)), "cgm_block_helmholtz")
        return blk

    def block_reference(self, k, row_shift=0.5, col_shift=0.0):
        t = self.torch
        blk = t.empty(self.N * self.N, dtype=t.complex128, device=self.device)
        check(lib().cgm_block_reference(self._ptr, float(k), float(row_shift), float(col_shift), _ptr(blk)),
              "cgm_block_reference")
        return blk

    # ---- operator ---------------------------------------------------------
    def matvec(self, gamma, chi, x, mode=_lib.OP_FORWARD, out=None):
        y = out if out is not None else self.torch.empty_like(x)
        check(lib().cgm_matvec(self._ptr, _ptr(gamma), _ptr(chi), _ptr(x), _ptr(y), int(mode)),
              "cgm_matvec")
        return y

    def matvec_batched(self, gamma, chi, x, mode=_lib.OP_FORWARD, out=None):
        """y[b] = A_b x[b] for x (B, N*N); chi (B, N*N) or (N*N,) shared."""
        y = out if out is not None else self.torch.empty_like(x)
        chi_stride = 0 if (chi is None or chi.dim() == 1) else self.N * self.N
        check(lib().cgm_matvec_batched(self._ptr, _ptr(gamma), _ptr(chi), chi_stride, _ptr(x), _ptr(y),
                                       int(x.shape[0]), int(mode)), "cgm_matvec_batched")
        return y

    def matvec_c128(self, gamma64, chi64, x, mode=_lib.OP_FORWARD, out=None):
        y = out if out is not None else self.torch.empty_like(x)
        check(lib().cgm_matvec_c128(self._ptr, _ptr(gamma64), _ptr(chi64), _ptr(x), _ptr(y), int(mode)),
              "cgm_matvec_c128")
        return y

    def synchronize(self):
        check(lib().cgm_synchronize(self._ptr), "cgm_synchronize")

    # ---- solvers ----------------------------------------------------------
    def gmres(self, gamma, chi, b, x=None, restart=30, tol=1e-6, max_cycles=100, adjoint=False,
              history=False):
        """Restarted GMRES(restart).  Returns (x, info dict).  x (if given) is
        the initial guess and is updated in place."""
        t = self.torch
        x0_zero = x is None
        if x is None:
            x = t.zeros_like(b)
        hist = None
        if history:
            hist = t.zeros(restart * max(max_cycles, 1) + 1, dtype=t.float64, device=self.device)
        p = _lib.GmresParams(int(restart), int(max_cycles), float(tol), int(bool(adjoint)), int(x0_zero))
        r = _lib.GmresResult()
        st = lib().cgm_gmres(self._ptr, _ptr(gamma), _ptr(chi), _ptr(b), _ptr(x), _c.byref(p),
                             _c.byref(r), _ptr(hist))
        check(st, "cgm_gmres", ok=(_lib.SUCCESS, _lib.ERROR_NOT_CONVERGED))
        info = dict(iterations=r.iterations, cycles=r.cycles, converged=bool(r.converged),
                    rel_residual=r.rel_residual, beta0=r.beta0, bnorm=r.bnorm)
        if history:
            info["history"] = hist[: r.iterations + 1].cpu().numpy()
        return x, info

    def set_batched_graphs(self, enable):
        """1 (default): batched restart cycles are captured once per call into a
        HIP graph and replayed; 0: launched one by one."""
        check(lib().cgm_set_batched_graphs(self._ptr, int(bool(enable))), "cgm_set_batched_graphs")

    def gmres_batched(self, gamma, chi, b, x=None, restart=30, tol=1e-6, max_cycles=100, history=False):
        """Lockstep GMRES(restart) of B systems: b (B, N*N) complex64, chi either
        (B, N*N) (own maps) or (N*N,) (one map, multiple right-hand sides).
        Returns (x, [info per system])."""
        t = self.torch
        B = b.shape[0]
        n = self.N * self.N
        chi_stride = 0 if chi.dim() == 1 else n
        x0_zero = x is None
        if x is None:
            x = t.zeros_like(b)
        hs = restart * max(max_cycles, 1) + 1
        hist = t.zeros(B, hs, dtype=t.float64, device=self.device) if history else None
        p = _lib.GmresParams(int(restart), int(max_cycles), float(tol), 0, int(x0_zero))
        res = (_lib.GmresResult * B)()
        st = lib().cgm_gmres_batched(self._ptr, _ptr(gamma), _ptr(chi), chi_stride, _ptr(b), _ptr(x), B,
                                     _c.byref(p), res, _ptr(hist))
        check(st, "cgm_gmres_batched", ok=(_lib.SUCCESS, _lib.ERROR_NOT_CONVERGED))
        infos = []
        h = hist.cpu().numpy() if history else None
        for k in range(B):
            r = res[k]
            d = dict(iterations=r.iterations, cycles=r.cycles, converged=bool(r.converged),
                     rel_residual=r.rel_residual, beta0=r.beta0, bnorm=r.bnorm)
            if history:
                d["history"] = h[k][: r.iterations + 1]
            infos.append(d)
        return x, infos

    def fgmres(self, gamma, chi, b, x=None, precond=None, restart=30, tol=1e-6, max_cycles=100,
               stall_ratio=0.5, history=False):
        """Right-preconditioned flexible GMRES(restart) with the preconditioner
        `precond`: the learned network (cugmres_amd.fno.Preconditioner), the
        tile inverses (cugmres_amd.tiles.Tiles), both (cugmres_amd.tiles.TwoLevel)
        or None (the identity).  A restart above 64 runs the long-recurrence
        path (the whole basis in device memory, up to 16384 steps per cycle).  Plain GMRES takes over when a preconditioned cycle leaves
        more than stall_ratio of the residual.  Returns (x, info dict)."""
        t = self.torch
        x0_zero = x is None
        if x is None:
            x = t.zeros_like(b)
        hist = None
        if history:
            hist = t.zeros(restart * max(max_cycles, 1) + 1, dtype=t.float64, device=self.device)
        p = _lib.FgmresParams(int(restart), int(max_cycles), float(tol), int(x0_zero), float(stall_ratio))
        r = _lib.FgmresResult()
        pc = _precond(precond)
        st = lib().cgm_fgmres_pc(self._ptr, _ptr(gamma), _ptr(chi), _ptr(b), _ptr(x), _c.byref(pc), _c.byref(p),
                                 _c.byref(r), _ptr(hist))
        check(st, "cgm_fgmres_pc", ok=(_lib.SUCCESS, _lib.ERROR_NOT_CONVERGED))
        info = dict(iterations=r.iterations, cycles=r.cycles, converged=bool(r.converged),
                    rel_residual=r.rel_residual, beta0=r.beta0, bnorm=r.bnorm,
                    precond_cycles=r.precond_cycles, fell_back=bool(r.fell_back))
        if history:
            info["history"] = hist[: r.iterations + 1].cpu().numpy()
        return x, info

    def fgmres_batched(self, gamma, chi, b, x=None, precond=None, restart=30, tol=1e-6, max_cycles=100,
                       history=False):
        """Lockstep flexible GMRES(restart) of B systems, one network forward
        pass per step for all of them: b (B, N*N) complex64, chi (B, N*N) or
        (N*N,) shared; precond a Preconditioner created with batch >= B (None
        = the identity).  No plain-GMRES fallback.  Returns (x, [info per
        system])."""
        t = self.torch
        B = b.shape[0]
        n = self.N * self.N
        chi_stride = 0 if chi.dim() == 1 else n
        x0_zero = x is None
        if x is None:
            x = t.zeros_like(b)
        hs = restart * max(max_cycles, 1) + 1
        hist = t.zeros(B, hs, dtype=t.float64, device=self.device) if history else None
        p = _lib.FgmresParams(int(restart), int(max_cycles), float(tol), int(x0_zero), 0.0)
        res = (_lib.FgmresResult * B)()
        pc = _precond(precond)
        st = lib().cgm_fgmres_pc_batched(self._ptr, _ptr(gamma), _ptr(chi), chi_stride, _ptr(b), _ptr(x),
                                         _c.byref(pc), B, _c.byref(p), res, _ptr(hist))
        check(st, "cgm_fgmres_pc_batched", ok=(_lib.SUCCESS, _lib.ERROR_NOT_CONVERGED))
        infos = []
        h = hist.cpu().numpy() if history else None
        for k in range(B):
            r = res[k]
            d = dict(iterations=r.iterations, cycles=r.cycles, converged=bool(r.converged),
                     rel_residual=r.rel_residual, beta0=r.beta0, bnorm=r.bnorm,
                     precond_cycles=r.precond_cycles, fell_back=bool(r.fell_back))
            if history:
                d["history"] = h[k][: r.iterations + 1]
            infos.append(d)
        return x, infos

    def gmres_refine(self, gamma32, gamma64, chi64, b, x=None, inner_restart=50, inner_tol=1e-4,
                     inner_max_cycles=40, outer_tol=1e-12, max_outer=30):
        t = self.torch
        if x is None:
            x = t.zeros_like(b)
        p = _lib.RefineParams(int(inner_restart), int(inner_max_cycles), float(inner_tol),
                              float(outer_tol), int(max_outer))
        r = _lib.RefineResult()
        st = lib().cgm_gmres_refine(self._ptr, _ptr(gamma32), _ptr(gamma64), _ptr(chi64), _ptr(b),
                                    _ptr(x), _c.byref(p), _c.byref(r))
        check(st, "cgm_gmres_refine", ok=(_lib.SUCCESS, _lib.ERROR_NOT_CONVERGED))
        return x, _refine_out(r)

    def gmres_refine_pc(self, gamma32, gamma64, chi64, b, precond, x=None, inner_restart=1000, inner_tol=1e-4,
                        inner_max_cycles=2, outer_tol=1e-12, max_outer=30):
        """FP64 refinement whose corrections are solved by flexible GMRES with
        `precond` (tiles.Tiles, tiles.TwoLevel or fno.Preconditioner; the tiles
        factored for chi64 rounded to float)."""
        t = self.torch
        if x is None:
            x = t.zeros_like(b)
        p = _lib.RefineParams(int(inner_restart), int(inner_max_cycles), float(inner_tol),
                              float(outer_tol), int(max_outer))
        r = _lib.RefineResult()
        pc = _precond(precond)
        st = lib().cgm_gmres_refine_pc(self._ptr, _ptr(gamma32), _ptr(gamma64), _ptr(chi64), _ptr(b), _ptr(x),
                                       _c.byref(pc), _c.byref(p), _c.byref(r))
        check(st, "cgm_gmres_refine_pc", ok=(_lib.SUCCESS, _lib.ERROR_NOT_CONVERGED))
        return x, _refine_out(r)


def _refine_out(r):
    return dict(outer=r.outer, inner_iterations=r.inner_iterations, converged=bool(r.converged),
                rel_residual=r.rel_residual)


def _precond(precond):
    """cgm_precond from a preconditioner object: the tile inverses
    (tiles.Tiles), the network (fno.Preconditioner) or None (the identity)."""
    pc = _lib.Precond(None, None)
    if precond is None:
        return pc
    from . import tiles as _tiles
    if isinstance(precond, _tiles.TwoLevel):
        pc.tiles = precond.tiles.handle
        pc.net = precond.net.handle
    elif isinstance(precond, _tiles.Tiles):
        pc.tiles = precond.handle
    else:
        pc.net = precond.handle
    return pc


def reference_chi(N):
    return float(lib().cgm_reference_chi(int(N)))


def n_timestamps(maxiter):
    return int(lib().cgm_n_timestamps(int(maxiter)))


def bessel01(x):
    out = [_c.c_double() for _ in range(4)]
    lib().cgm_bessel01(float(x), *[_c.byref(v) for v in out])
    return tuple(v.value for v in out)
