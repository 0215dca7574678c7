"""Seeded problem generators for the benchmark workloads (host side, numpy).

The reference builds one fixed problem (an eps = 2.25 disk, x0 = b, plane wave
along the row index; B/CUDA_GMRES/GMRES_kernels.cuh:12-19,557-572).  The
benchmark workloads need seeded permittivity maps and incidence angles that
are identical wherever they are generated, so the random stream is this
module's own SplitMix64 (Steele, Lea & Flood, OOPSLA 2014): a 64-bit counter
advanced by 0x9E3779B97F4A7C15 and finalised by the two xor-shift-multiply
rounds below.  Uniforms are the top 53 bits times 2**-53.

Orientation: arrays are (N, N) row-major, element [i, j] is row i (the
direction along which the reference's plane wave travels), column j.
"""
import math

import numpy as np

_GOLDEN = 0x9E3779B97F4A7C15
_M64 = (1 << 64) - 1

EPS_HI = 12.1          # high permittivity of the binary maps
WAVELENGTHS_PER_CELL = 1.0 / 20.0   # h = lambda / 20
KH = 2.0 * math.pi * WAVELENGTHS_PER_CELL   # k0 * h


class SplitMix64:
    def __init__(self, seed):
        self.state = int(seed) & _M64

    def next_u64(self):
        self.state = (self.state + _GOLDEN) & _M64
        z = self.state
        z = ((z ^ (z >> 30)) * 0xBF58476D1CE4E5B9) & _M64
        z = ((z ^ (z >> 27)) * 0x94D049BB133111EB) & _M64
        return z ^ (z >> 31)

    def uniform(self):
        return (self.next_u64() >> 11) * (1.0 / (1 << 53))

    def normal(self):
        """Box-Muller, one value per call (the sine partner is discarded)."""
        u1 = self.uniform()
        u2 = self.uniform()
        u1 = max(u1, 1e-300)
        return math.sqrt(-2.0 * math.log(u1)) * math.cos(2.0 * math.pi * u2)


def block_eps_map(N, block, p=0.5, fill=1.0, seed=0, eps_hi=EPS_HI, rng=None):
    """Binary permittivity map of `block` x `block` cells with values {1, eps_hi}.

    The random blocks cover a centred square of nb x nb blocks,
    nb = round(sqrt(fill) * N / block) (fill is the area fraction); cells
    outside it are eps = 1.  Blocks are drawn in row-major order, a block is
    eps_hi when its uniform draw is < p.
    """
    rng = rng if rng is not None else SplitMix64(seed)
    nb = int(round(math.sqrt(fill) * N / block))
    nb = max(0, min(nb, N // block))
    off = (N - nb * block) // 2
    eps = np.ones((N, N), np.float64)
    for bi in range(nb):
        for bj in range(nb):
            if rng.uniform() < p:
                i0, j0 = off + bi * block, off + bj * block
                eps[i0:i0 + block, j0:j0 + block] = eps_hi
    return eps


def smooth_eps_map(N, sigma=4.0, seed=0, lo=1.0, hi=EPS_HI):
    """White noise N(0,1) per cell, periodic Gaussian filter of width sigma
    cells (applied by FFT), then min-max mapped to [lo, hi]."""
    rng = SplitMix64(seed)
    noise = np.array([rng.normal() for _ in range(N * N)], np.float64).reshape(N, N)
    f = np.fft.fftfreq(N)
    g1 = np.exp(-2.0 * (math.pi * sigma * f) ** 2)
    sm = np.real(np.fft.ifft2(np.fft.fft2(noise) * np.outer(g1, g1)))
    a, b = sm.min(), sm.max()
    return lo + (hi - lo) * (sm - a) / (b - a)


def plane_wave(N, kh=KH, theta=0.0):
    """b[i, j] = exp(-i kh (i cos theta + j sin theta)), the reference's sign
    convention (B/CUDA_GMRES/GMRES_kernels.cuh:197-198)."""
    i = np.arange(N, dtype=np.float64)[:, None]
    j = np.arange(N, dtype=np.float64)[None, :]
    return np.exp(-1j * kh * (i * math.cos(theta) + j * math.sin(theta)))


def config_plain(N=512, seed=0):
    """Plain GMRES workload: 8x8 blocks, p = 0.5, central 75 % of the area,
    plane wave at a seeded angle in [0, 2 pi).  Returns (chi, b, theta) with
    chi = eps - 1 (the operator is A = I - (k0^2 G) diag(chi))."""
    rng = SplitMix64(seed)
    theta = 2.0 * math.pi * rng.uniform()
    eps = block_eps_map(N, 8, 0.5, 0.75, rng=rng)
    return eps - 1.0, plane_wave(N, KH, theta), theta


def config_cpu(N=64, seed=0):
    """CPU workload: 4x4 blocks over the whole grid, x-directed plane wave."""
    eps = block_eps_map(N, 4, 0.5, 1.0, seed=seed)
    return eps - 1.0, plane_wave(N, KH, 0.0), 0.0


def config_mixed(N=1024, seed=0):
    """Mixed-precision workload: smooth map in [1, 12.1], sigma = 4 cells."""
    eps = smooth_eps_map(N, 4.0, seed)
    return eps - 1.0, plane_wave(N, KH, 0.0), 0.0
