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
acts on each tile alone (a vector supported on one tile maps into that tile);
the reduced form (tile_reduced) applies the same P as the dense inverses.
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


def tile_reduced(blk, chi, N, T):
    """The reduced form of every tile's inverse.  With S the cells of the tile
    where chi != 0, D_S = diag(chi_S) and G_SS the rows and columns S of G_t,

        A_t^{-1} = I + G_t[:, S] W_t R_S,   W_t = (I - D_S G_SS)^{-1} D_S,

    (R_S restricts a tile vector to S): the columns of A_t outside S are unit
    vectors, so the inverse differs from I only through G_t acting on a vector
    supported on S.  Returns per tile (a, b): (S, W_t)."""
    assert N % T == 0
    blk = np.asarray(blk, np.complex128)
    c = to_tiles(np.asarray(chi, np.float64).reshape(-1), N, T)
    out = {}
    for a in range(N // T):
        for b in range(N // T):
            S = np.flatnonzero(c[a, b] != 0.0)
            d = c[a, b][S]
            M = np.eye(S.size) - d[:, None] * _green(blk, T, S, S)
            out[a, b] = (S, np.linalg.solve(M, np.diag(d)) if S.size else np.zeros((0, 0)))
    return out


def _green(blk, T, P, Q):
    """Rows P, columns Q (tile cells, row-major) of the tile's Green's matrix."""
    pi, pj = np.divmod(np.asarray(P), T)
    qi, qj = np.divmod(np.asarray(Q), T)
    return blk[np.abs(pi[:, None] - qi[None, :]), np.abs(pj[:, None] - qj[None, :])]


def tile_apply_reduced(R, blk, v, N, T):
    """P v from tile_reduced's (S, W) per tile: v_t + G_t[:, S] (W v_t[S])."""
    blk = np.asarray(blk, np.complex128)
    vt = to_tiles(np.asarray(v, np.complex128).reshape(-1), N, T)
    yt = vt.copy()
    for (a, b), (S, W) in R.items():
        if S.size:
            u = W @ vt[a, b][S]
            for p0 in range(0, T * T, 2048):   # G[:, S] u in row blocks
                yt[a, b][p0:p0 + 2048] += _green(blk, T, np.arange(p0, min(T * T, p0 + 2048)), S) @ u
    return from_tiles(yt, N, T)
