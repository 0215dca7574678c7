"""The reference Python module's functions (B/PyGMRES/GMRES.py) over the C ABI.

Same names, argument meaning and return shapes as the reference module, in
its orientation (the transpose of the C++ one: the plane wave runs along the
column index and the half-cell shift of the reference operator is on the
column offset).  Host arrays in, host arrays out; the operator work runs on
the GPU through libcugmres.so.  Importing this module runs nothing (the
reference module runs its whole benchmark at import, B/PyGMRES/GMRES.py:95-221).

    get_plane_wave(k, size)                  B/PyGMRES/GMRES.py:13-18
    get_greenfun(r, k)                       B/PyGMRES/GMRES.py:20-21
    get_green_matrix(k, size)                B/PyGMRES/GMRES.py:23-27
    G_matvec(vec, k, qweq=False, g=None)     B/PyGMRES/GMRES.py:30-44
    matvec(x, eps, k, qweq=False, transpose=False, g=None)   B/PyGMRES/GMRES.py:47-57
    get_eps_from_mask(e, mask)               B/PyGMRES/GMRES.py:60-61
    get_toeplitz_mat(ij_block)               B/PyGMRES/GMRES.py:64-69
    get_complex_array(filename)              B/PyGMRES/GMRES.py:71-92
    gmres(x0, eps, k, ...)                   the loop of B/PyGMRES/GMRES.py:95-221 as a function

Arithmetic: complex64 operator on the device (the reference mixes complex64
and complex128); results are returned as complex128 column vectors like
the reference's.  The spectrum of the kernel is computed once per (size, k)
or per distinct `g` array and cached.
"""
import ctypes

import numpy as np

from . import formats

_cache = {}


def _ctx(N):
    from . import solver
    key = ("ctx", N)
    if key not in _cache:
        _cache[key] = solver.Context(N)
    return _cache[key]


def get_plane_wave(k, size):
    """exp(-1j k x) on a size x size grid, x the column index: the reference's
    wave at incidence angle 0 (row r holds the same phase ramp for every r)."""
    ramp = np.exp(-1j * k * np.arange(size, dtype=np.float64))
    return np.broadcast_to(ramp, (size, size)).copy()


def get_greenfun(r, k):
    """(1j/4) H0^(1)(k r), FP64 (own Bessel functions, host)."""
    from ._lib import lib
    r = np.asarray(r, dtype=np.float64)
    flat = np.ascontiguousarray(r.reshape(-1))
    out = np.empty(flat.size, np.complex128)
    lib().cgm_green_values(flat.ctypes.data_as(ctypes.c_void_p), flat.size, float(k),
                           out.ctypes.data_as(ctypes.c_void_p))
    return out.reshape(r.shape)


def get_green_matrix(k, size):
    """Kernel block g[r, c] = (i/4) H0(k |(r, c - 1/2)|): distances with the
    half-cell shift on the column offset (the Python orientation)."""
    rows = np.arange(size, dtype=np.float64)[:, None]
    cols = np.arange(size, dtype=np.float64)[None, :] - 0.5
    return get_greenfun(np.hypot(rows, cols), k)


def get_toeplitz_mat(ij_block):
    """(2N-1)^2 even circulant embedding of an N x N block: entry (r, c) is
    block[d(r), d(c)] with d(t) = t for t < N and 2N - 1 - t above (the
    mirrored negative offsets)."""
    blk = np.asarray(ij_block)
    n = blk.shape[0]
    t = np.arange(2 * n - 1)
    d = np.where(t < n, t, 2 * n - 1 - t)
    return blk[d[:, None], d[None, :]].copy()


def get_eps_from_mask(e, mask):
    """Per-cell contrast column (e - 1) * mask, shape (N^2, 1)."""
    return np.multiply(np.asarray(mask).reshape(-1, 1), e - 1)


def get_complex_array(filename):
    return formats.read_complex(filename, np.complex64)


def _spectrum(N, k, g):
    if g is None:
        key = ("spec", N, float(k))
        if key not in _cache:
            # Python orientation: half-cell shift on the column offset
            _cache[key] = _ctx(N).spectrum_reference(float(k), 0.0, 0.5)
        return _cache[key]
    g = np.ascontiguousarray(g, np.complex64)
    key = ("specg", N, g.tobytes().__hash__())
    if key not in _cache:
        _cache[key] = _ctx(N).spectrum_from_gprep(g)
    return _cache[key]


def G_matvec(vec, k, qweq=False, g=None):
    """G vec for the embedded kernel; returns an (N^2, 1) column.  With
    qweq=True returns (G_block, G) like the reference."""
    vec = np.asarray(vec)
    size = int(np.sqrt(vec.shape[0]))
    if qweq:
        if g is None:
            G_block = get_green_matrix(k, size)
            return G_block, get_toeplitz_mat(G_block)
        return None, g
    from . import _lib
    c = _ctx(size)
    gam = _spectrum(size, k, g)
    x = c.to_device(vec.reshape(-1), "c64")
    y = c.matvec(gam, None, x, _lib.OP_GREEN)
    return y.cpu().numpy().astype(np.complex128).reshape((-1, 1))


def matvec(x, eps, k, qweq=False, transpose=False, g=None):
    """x - G(chi x) (or x - chi G(x) when transpose) with chi = k^2 eps."""
    x = np.asarray(x).reshape((-1, 1))
    chi = (k ** 2) * np.asarray(eps).reshape((-1, 1))
    if qweq:
        return G_matvec(x * chi, k, qweq, g)
    from . import _lib
    size = int(np.sqrt(x.shape[0]))
    c = _ctx(size)
    gam = _spectrum(size, k, g)
    dx = c.to_device(x.reshape(-1), "c64")
    dchi = c.to_device(np.real(chi).reshape(-1), "f32")
    y = c.matvec(gam, dchi, dx, _lib.OP_ADJOINT if transpose else _lib.OP_FORWARD)
    return y.cpu().numpy().astype(np.complex128).reshape((-1, 1))


def gmres(x0, eps, k, b=None, restart=30, tol=1e-6, max_cycles=100, transpose=False, g=None,
          history=False):
    """Solve matvec(x) = b (b defaults to x0, the reference's choice x0 = b,
    B/PyGMRES/GMRES.py:112) with restarted GMRES on the device.  Returns
    (x as an (N^2, 1) column, info dict)."""
    x0 = np.asarray(x0).reshape(-1)
    size = int(np.sqrt(x0.size))
    c = _ctx(size)
    gam = _spectrum(size, k, g)
    chi = np.real((k ** 2) * np.asarray(eps).reshape(-1))
    rhs = x0 if b is None else np.asarray(b).reshape(-1)
    dx = c.to_device(x0, "c64")
    x, info = c.gmres(gam, c.to_device(chi, "f32"), c.to_device(rhs, "c64"), dx, restart=restart,
                      tol=tol, max_cycles=max_cycles, adjoint=transpose, history=history)
    return x.cpu().numpy().astype(np.complex128).reshape((-1, 1)), info
