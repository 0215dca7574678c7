"""Analytical field of a plane wave scattered by a homogeneous dielectric disk
(the Mie series of the 2-D scalar Helmholtz problem).

The reference ships such fields as data (B/CUDA_GMRES/IN/analytical_solution_N.txt,
read at B/CUDA_GMRES/launch_GMRES.cuh:26-78 and compared at 195-218) but not the
code that made them; this module computes them from the textbook series
(Bohren & Huffman, "Absorption and Scattering of Light by Small Particles",
1983, sec. 8.4; the scalar/TM case).

Conventions are the integral equation's: the Green's function is the
outgoing (i/4) H0^(1)(k r), so the scattered field is a sum of Hankel
functions of the first kind.  With the disk of radius a and relative
permittivity eps centred at (ci, cj), inside wavenumber k1 = k sqrt(eps),
polar angle theta of (i - ci, j - cj) and the incident wave
exp(i k r cos(theta - theta_inc)) = sum_n i^n J_n(k r) e^{i n (theta - theta_inc)}
(phase zero at the centre):

    r > a:  u = e^{i k r cos(theta - theta_inc)} + sum_n i^n b_n H_n(k r) e^{i n (theta - theta_inc)}
    r < a:  u = sum_n i^n  c_n J_n(k1 r)            e^{i n (theta - theta_inc)}
    b_n = (k1 J_n'(k1 a) J_n(k a) - k J_n(k1 a) J_n'(k a))
        / (k J_n(k1 a) H_n'(k a) - k1 J_n'(k1 a) H_n(k a))
    c_n = (J_n(k a) + b_n H_n(k a)) / J_n(k1 a)

from continuity of u and du/dr at r = a.  The reference's incident wave
exp(-i k i) (B/CUDA_GMRES/GMRES_kernels.cuh:12-19,557-572) travels towards
-i: theta_inc = pi and the phase factor exp(-i k ci) moves its zero to row 0.
"""
import math

import numpy as np


def series_terms(k, eps, a):
    """Orders kept: |n| <= k1 a + 4 (k1 a)^(1/3) + 10 (Wiscombe's criterion)."""
    x = k * math.sqrt(eps) * a
    return int(x + 4.0 * x ** (1.0 / 3.0) + 10)


def coefficients(k, eps, a, nmax=None):
    """(n, b_n, c_n) for n = -nmax..nmax."""
    from scipy.special import h1vp, hankel1, jv, jvp
    nmax = series_terms(k, eps, a) if nmax is None else nmax
    n = np.arange(-nmax, nmax + 1)
    k1 = k * math.sqrt(eps)
    ka, k1a = k * a, k1 * a
    num = k1 * jvp(n, k1a) * jv(n, ka) - k * jv(n, k1a) * jvp(n, ka)
    den = k * jv(n, k1a) * h1vp(n, ka) - k1 * jvp(n, k1a) * hankel1(n, ka)
    b = num / den
    c = (jv(n, ka) + b * hankel1(n, ka)) / jv(n, k1a)
    return n, b, c


def disk_field(pi_, pj, k, eps, ci, cj, a, theta_inc=math.pi, phase_origin=None):
    """Total field at the points (pi_, pj) (arrays of equal shape, cell units).
    phase_origin = (i0, j0) puts the incident wave's phase zero at that point
    instead of the disk centre."""
    from scipy.special import hankel1, jv
    pi_ = np.asarray(pi_, np.float64)
    pj = np.asarray(pj, np.float64)
    x, y = pi_ - ci, pj - cj
    r = np.hypot(x, y)
    th = np.arctan2(y, x)
    inside = r < a
    k1 = k * math.sqrt(eps)
    n, b, c = coefficients(k, eps, a)
    # outside: the incident wave in closed form plus the scattered series
    # (whose coefficients decay past |n| ~ k1 a); inside: the series
    u = np.where(inside, 0.0, np.exp(1j * k * r * np.cos(th - theta_inc))).astype(np.complex128)
    ro = np.where(inside, a, r)       # keep H_n away from r = 0 where unused
    ri = np.where(inside, r, 0.0)
    for nn, bn, cn in zip(n, b, c):
        ph = (1j) ** int(nn % 4) * np.exp(1j * nn * (th - theta_inc))
        out = bn * hankel1(nn, k * ro)
        inn = cn * jv(nn, k1 * ri)
        u += ph * np.where(inside, inn, out)
    if phase_origin is not None:
        # exp(i k (p - c).d) with d the direction of travel: shift the origin
        d = (math.cos(theta_inc), math.sin(theta_inc))
        u *= np.exp(1j * k * ((ci - phase_origin[0]) * d[0] + (cj - phase_origin[1]) * d[1]))
    return u


def reference_problem_field(N, stride=1, k=None):
    """The reference's benchmark problem (eps = 2.25 disk of radius N/6 at
    (N/3, N/2), k = 2*pi*6/N per cell unless given, incident exp(-i k i)) on
    the N x N cell centres, every `stride`-th row and column.  Returns
    (field, rows, cols)."""
    k = 2.0 * math.pi / (N / 6.0) if k is None else float(k)
    rows = np.arange(0, N, stride)
    cols = np.arange(0, N, stride)
    I, J = np.meshgrid(rows, cols, indexing="ij")
    u = disk_field(I, J, k, 2.25, N / 3.0, N / 2.0, N / 6.0, math.pi, phase_origin=(0.0, 0.0))
    return u, rows, cols
