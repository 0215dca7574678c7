"""Tile (block-Jacobi) preconditioner: exact inverses of the operator on
T x T tiles of the grid (cgm_tiles_* in include/cugmres.h).

The reference has no preconditioner (B/CUDA_GMRES/GMRES.cuh:22-446).  On
high-contrast maps every high-permittivity block is a near-resonant
scatterer and unpreconditioned GMRES needs a number of steps that grows about
as N^1.8; the tile inverses remove those local resonances.  The factorisation
runs once per map (FP32 complex batched LU on the device), the application
once per Krylov step (one work-group per tile).
"""
import ctypes

from . import _lib
from ._lib import check, lib


class Tiles:
    """Inverses of the tile operators for up to `maps` contrast maps."""

    def __init__(self, ctx, T=16, maps=1):
        self.ctx = ctx
        self.T = int(T)
        self.maps_capacity = int(maps)
        h = ctypes.c_void_p()
        check(lib().cgm_tiles_create(ctx.handle, self.T, self.maps_capacity, ctypes.byref(h)), "cgm_tiles_create")
        self._h = h
        self.maps = 0

    @property
    def handle(self):
        return self._h

    def memory_bytes(self):
        b = ctypes.c_size_t()
        check(lib().cgm_tiles_memory(self._h, ctypes.byref(b)), "cgm_tiles_memory")
        return b.value

    def apply_bytes(self, batch=1):
        """Algorithmic bytes of one application to `batch` systems (after factor)."""
        b = ctypes.c_double()
        check(lib().cgm_tiles_apply_bytes(self._h, int(batch), ctypes.byref(b)), "cgm_tiles_apply_bytes")
        return b.value

    def factor(self, block, chi):
        """Invert every tile's operator: block is the N x N complex128 generator
        block on the device (Context.block_helmholtz), chi is (N*N,) for one map
        or (maps, N*N) float32 on the device."""
        maps = 1 if chi.dim() == 1 else int(chi.shape[0])
        stride = 0 if maps == 1 else self.ctx.N * self.ctx.N
        check(lib().cgm_tiles_factor(self._h, ctypes.c_void_p(block.data_ptr()), ctypes.c_void_p(chi.data_ptr()),
                                     stride, maps), "cgm_tiles_factor")
        self.maps = maps
        return self

    def apply(self, x, out=None):
        """y = P x for x (N*N,) or (B, N*N) complex64 on the device."""
        y = out if out is not None else self.ctx.torch.empty_like(x)
        B = 1 if x.dim() == 1 else int(x.shape[0])
        check(lib().cgm_tiles_apply(self._h, ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()), B),
              "cgm_tiles_apply")
        return y

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().cgm_tiles_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class TwoLevel:
    """The tile inverses followed by the learned network as a correction on
    what they leave (cgm_precond with both members set):

        z1 = T v,   r = v - A z1,   P(v) = z1 + ||r|| F(r / ||r||)

    The network F is trained on the normalised residuals r / ||r|| of the tile
    solves (cugmres_amd.learn.train with tiles_T)."""

    def __init__(self, tiles, net):
        self.tiles = tiles
        self.net = net

    def apply(self, ctx, gamma, chi, v):
        """P(v) for unit vectors v (B, N*N) on the device: the composition the
        library applies, written with the library's own calls (for training)."""
        z1, r, rho = tile_residual(ctx, gamma, self.tiles, chi, v)
        z2 = self.net.forward(chi, r)
        return z1 + rho[:, None] * z2


def tile_residual(ctx, gamma, tiles, chi, v):
    """(z1, r / ||r||, ||r||) with z1 = T v and r = v - A z1, per row of v."""
    from ._lib import OP_FORWARD
    t = ctx.torch
    z1 = tiles.apply(v)
    az = ctx.matvec_batched(gamma, chi, z1, OP_FORWARD)
    r = v - az
    rho = t.linalg.vector_norm(r, dim=1)
    return z1, r / t.clamp(rho, min=1e-30)[:, None], rho
