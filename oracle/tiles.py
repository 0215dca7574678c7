"""Tile (block-Jacobi) preconditioner of the GFIEM operator, FP64 numpy.

TEST INFRASTRUCTURE ONLY: the checker of cgm_tiles_* (include/cugmres.h).
The reference has no preconditioner (B/CUDA_GMRES/GMRES.cuh:22-446 runs plain
GMRES); this restates the north-star extension the library adds.

The N x N grid is cut into square tiles of T x T cells.  On tile t the
operator restricted to the tile is

    A_t = I - G_t diag(chi_t),   G_t[p, q] = g[|p_i - q_i|, |p_j - q_j|],

with g the N x N generator block of the operator (the same block whose
circulant embedding gives the kernel spectrum), p, q cells of the tile in
row-major order.  The preconditioner applies the exact inverse of every
tile's operator to that tile's part of the vector:

    (P v)_t = A_t^{-1} v_t.

Pins (tests/test_tiles.py): a single tile covering the grid makes P = A^{-1}
(then A P v = v, checked with the oracle's FFT matvec, which is itself pinned
by the reference's MKL dumps and dense brute force); chi = 0 makes P = I; P
acts on each tile alone (a vector supported on one tile maps into that tile).
"""
import numpy as np


def tile_green(blk, T):
    """Dense T^2 x T^2 Green's matrix of one tile from the generator block
    blk[|di|, |dj|] (complex128, at least T x T)."""
    i, j = np.divmod(np.arange(T * T), T)
    return blk[np.abs(i[:, None] - i[None, :]), np.abs(j[:, None] - j[None, :])]


def to_tiles(v, N, T):
    """(N*N,) -> (N/T, N/T, T*T): tile (a, b) holds cells (a*T + r, b*T + c)."""
    return v.reshape(N // T, T, N // T, T).transpose(0, 2, 1, 3).reshape(N // T, N // T, T * T)


def from_tiles(t, N, T):
    return t.reshape(N // T, N // T, T, T).transpose(0, 2, 1, 3).reshape(N * N)


def tile_inverses(blk, chi, N, T):
    """A_t^{-1} of every tile, shape (N/T, N/T, T^2, T^2), complex128."""
    assert N % T == 0
    G = tile_green(np.asarray(blk, np.complex128), T)
    c = to_tiles(np.asarray(chi, np.float64).reshape(-1), N, T)
    A = np.eye(T * T)[None, None] - G[None, None] * c[:, :, None, :]
    return np.linalg.inv(A)


def tile_apply(K, v, N, T):
    """P v for the inverses K of tile_inverses."""
    vt = to_tiles(np.asarray(v, np.complex128).reshape(-1), N, T)
    return from_tiles(np.einsum("abpq,abq->abp", K, vt), N, T)
