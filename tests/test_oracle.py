"""Pins of the CPU oracle against the reference's own fixtures, closed forms
and textbook routines (all CPU)."""
import hashlib
import math

import numpy as np
import pytest

from conftest import golden, rel
from cugmres_amd import formats, problems


# ---- special functions --------------------------------------------------------
def test_bessel_vs_scipy_fixture(orc):
    rows = golden("hankel_scipy.json")["rows"]
    for x, j0, j1, y0, y1 in rows:
        got = orc.bessel01(x)
        env = max(1.0, math.sqrt(2 / (math.pi * x))) if x > 1 else 1.0
        for g, r in zip(got, (j0, j1, y0, y1)):
            scale = max(abs(r), 1e-300) if x <= 1 else env
            assert abs(g - r) <= 2e-13 * max(scale, abs(r)), (x, g, r)


def test_bessel_wronskian(orc):
    for x in np.logspace(-2, math.log10(480.0), 300):
        j0, j1, y0, y1 = orc.bessel01(x)
        assert abs((j1 * y0 - j0 * y1) - 2 / (math.pi * x)) < 1e-13 * max(1.0, 2 / (math.pi * x))


# ---- FFT --------------------------------------------------------------------------
@pytest.mark.parametrize("n", [1, 2, 3, 5, 7, 12, 96, 127, 384, 1023, 1024])
def test_fft_vs_numpy(orc, n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    assert rel(orc.fft(x, -1), np.fft.fft(x)) < 1e-14
    assert rel(orc.fft(x, +1), np.fft.ifft(x) * n) < 1e-14


def test_fft2_vs_numpy(orc):
    rng = np.random.default_rng(3)
    x = rng.standard_normal((48, 48)) + 1j * rng.standard_normal((48, 48))
    assert rel(orc.fft2(x), np.fft.fft2(x)) < 1e-14


# ---- reference fixtures -----------------------------------------------------------
@pytest.mark.parametrize("N", [32, 64, 128])
def test_gprep_matches_reference(orc, N):
    g = golden("gprep.npz")
    mine = orc.ref_gprep(N).reshape(-1)
    if N == 128:
        ref = g["g128_val"]
        mine = mine[g["g128_idx"]]
    else:
        ref = g[f"g{N}"]
    # float rounding of values |G| < 0.41 printed with 12 decimals
    assert np.abs(mine - ref).max() < 2e-7
    blk, sym = orc.gprep_to_block(N, orc.ref_gprep(N))
    assert sym


@pytest.mark.parametrize("N", [32, 64, 128, 256, 512, 1024])
def test_cylinder_mask_bit_exact(orc, N):
    m = orc.ref_cylinder_mask(N)
    info = golden("masks.json")[str(N)]
    assert int(m.sum()) == info["ones"]
    assert hashlib.sha256(formats.mask_text(m).encode()).hexdigest() == info["sha256"]


def test_reference_chi_and_wave_number(orc):
    # CHI = 1.25 k^2 with k = 6.28/(N/6), evaluated in float
    for N in [32, 256, 512]:
        k = 6.28 / (N / 6)
        assert abs(orc.ref_chi(N) - 1.25 * k * k) < 1e-6 * 1.25 * k * k
        assert abs(orc.ref_wave_number(N) - k) < 1e-6 * k


# ---- operator ---------------------------------------------------------------------
@pytest.mark.parametrize("N,M", [(6, 11), (6, 12), (9, 20), (12, 24)])
def test_fft_matvec_equals_dense_bttb(orc, N, M):
    rng = np.random.default_rng(N + M)
    blk = orc.baseline_green_block(N, 2 * math.pi / 20)
    chi = rng.uniform(0, 11.1, N * N)
    x = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
    op = orc.Op(N, blk, M)
    for mode in (0, 1, 2):
        assert rel(op.matvec(chi, x, mode), orc.dense_matvec(N, blk, chi, x, mode)) < 1e-13


def test_identity_when_eps_is_one(orc):
    N = 16
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH))
    x = np.random.default_rng(1).standard_normal(N * N) + 0j
    assert rel(op.matvec(np.zeros(N * N), x), x) < 1e-15


def test_adjoint_identity(orc):
    """<A x, y> = <x, A^T y> (bilinear): G is complex symmetric."""
    N = 20
    rng = np.random.default_rng(5)
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH))
    chi = rng.uniform(0, 11.1, N * N)
    x = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
    y = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
    lhs = np.sum(op.matvec(chi, x, 0) * y)
    rhs = np.sum(x * op.matvec(chi, y, 1))
    assert abs(lhs - rhs) < 1e-12 * abs(lhs)


def test_self_term_small_ka_limit(orc):
    """Small-ka series of (i pi ka/2) H1(ka) - 1 (A&S 9.1.10-11):
    real part -(ka^2/2)(ln(ka/2) + gamma - 1/2), imaginary part pi ka^2/4."""
    kh = 2 * math.pi / 400
    blk = orc.baseline_green_block(4, kh)
    ka = kh / math.sqrt(math.pi)
    re_expect = -(ka * ka / 2) * (math.log(ka / 2) + 0.5772156649015329 - 0.5)
    im_expect = math.pi * ka * ka / 4
    assert abs(blk[0, 0].real - re_expect) < 1e-3 * abs(re_expect)
    assert abs(blk[0, 0].imag - im_expect) < 1e-3 * im_expect


# ---- MKL dumps (reference-mode operator at N = 256) --------------------------------
@pytest.fixture(scope="module")
def ref256(orc):
    N = 256
    blk, _ = orc.gprep_to_block(N, orc.ref_gprep(N))
    op = orc.Op(N, blk)
    chi = orc.ref_chi(N) * orc.ref_cylinder_mask(N).reshape(-1).astype(np.float64)
    b = orc.ref_plane_wave(N).reshape(-1).astype(np.complex128)
    return op, chi, b


def _cmp_strided(d, key, full):
    idx = d["idx"]
    assert rel(full[idx], d[key]) < 5e-6
    assert abs(np.linalg.norm(full) - d[key + "_norm"]) < 5e-6 * d[key + "_norm"]


def test_mkl_dumps_r0_AV0_V1_AV2(orc, ref256):
    op, chi, b = ref256
    d = golden("mkl_dumps.npz")
    r0 = b - op.matvec(chi, b)
    _cmp_strided(d, "r0", r0)
    _cmp_strided(d, "r0b", r0)
    _, st = op.gmres_reference(chi, b, b, 3)
    V = st["V"]
    _cmp_strided(d, "AV0", op.matvec(chi, V[0]))
    _cmp_strided(d, "V1", V[1])
    _cmp_strided(d, "AV2", op.matvec(chi, V[2]))
    # H after two steps (columns 0, 1 of the 4 x 3 array)
    H = st["H"]
    Hd = d["H"].reshape(4, 3)
    for (r, c) in [(0, 0), (0, 1), (1, 0), (1, 1), (2, 1)]:
        assert abs(H[r, c] - Hd[r, c]) < 1e-5 * max(1.0, abs(Hd[r, c]))  # complex64 run
    # Jtotal after the first rotation: [[c, s], [-conj(s), c]], identity elsewhere
    _, st1 = op.gmres_reference(chi, b, b, 1)
    J1 = st1["J"]
    Jd = d["Jtotal"].reshape(4, 4)
    assert np.abs(J1 - Jd[:2, :2]).max() < 2e-6
    assert np.abs(Jd[2:, 2:] - np.eye(2)).max() == 0
    # Htemp = Jtotal(after step 0) * H[:, 0:2]  (3 x 2)
    J3 = np.eye(3, dtype=np.complex128)
    J3[:2, :2] = J1
    Ht = J3 @ H[:3, :2]
    assert np.abs(Ht.reshape(-1) - d["Htemp"]).max() < 1e-5 * np.abs(d["Htemp"]).max()


@pytest.mark.slow
def test_reference_solve_matches_mkl_solution(orc, ref256):
    op, chi, b = ref256
    g = golden("mkl_solution.npz")
    x, st = op.gmres_reference(chi, b, b, 44)
    # FP64 run of the reference algorithm vs the reference's complex64 MKL run
    assert rel(g["solution"], x) < 1.2e-5
    # physics: relative error vs the Mie field, 0.238 at N = 256
    assert abs(rel(x, g["analytical"]) - 0.23810) < 2e-4
    # the converged FP64 solution is the same to 1e-11
    xc, info = op.gmres(chi, b, x0=b, restart=100, tol=1e-12, max_cycles=3)
    assert info["converged"] and rel(x, xc) < 1e-10


def test_python_history_prefix(orc):
    """First 10 logged residuals of B/PyGMRES/GMRES.py (Python orientation,
    k in FP64, |c[GMRES_i]| after GMRES_i + 1 steps)."""
    N = 256
    k = 2 * 3.14 / (N / 6)
    op = orc.Op(N, orc.shifted_green_block(N, k, 0.0, 0.5))
    x, y = np.mgrid[0:N, 0:N]
    mask = ((y - N / 3) ** 2 + (x - N / 2) ** 2 <= (N / 6) ** 2).astype(np.float64)
    chi = (k * k * 1.25 * mask).reshape(-1)
    b = np.exp(-1j * k * np.mgrid[:N, :N][1]).reshape(-1)
    _, st = op.gmres_reference(chi, b, b, 16)
    ref = golden("python_histories.npz")["residuals_256"]
    mine = st["res"][0] * np.abs(st["J"][4:14, 0])
    assert np.max(np.abs(mine - ref[:10]) / ref[:10]) < 1e-5


# ---- solvers ----------------------------------------------------------------------
def test_gmres_vs_dense_solve(orc):
    N = 8
    rng = np.random.default_rng(11)
    blk = orc.baseline_green_block(N, problems.KH)
    chi = rng.uniform(0, 3.0, N * N)
    A = np.stack([orc.dense_matvec(N, blk, chi, e, 0) for e in np.eye(N * N)], axis=1)
    b = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
    xd = np.linalg.solve(A, b)
    op = orc.Op(N, blk)
    x, info = op.gmres(chi, b, restart=64, tol=1e-13, max_cycles=3)
    assert info["converged"] and rel(x, xd) < 1e-11
    xa, info = op.gmres(chi, b, restart=64, tol=1e-13, max_cycles=3, adjoint=True)
    assert rel(xa, np.linalg.solve(A.T, b)) < 1e-11


def test_restart_equivalence(orc):
    """GMRES(m) with m >= iterations needed is full GMRES: same iterate."""
    N = 16
    rng = np.random.default_rng(2)
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH))
    chi = rng.uniform(0, 1.25, N * N)
    b = problems.plane_wave(N).reshape(-1)
    x1, i1 = op.gmres(chi, b, restart=200, tol=1e-10, max_cycles=1)
    x2, i2 = op.gmres(chi, b, restart=i1["iterations"], tol=1e-10, max_cycles=1)
    assert i1["converged"] and rel(x2, x1) < 1e-12


def test_refinement_reaches_fp64(orc):
    N = 24
    rng = np.random.default_rng(7)
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH))
    chi = rng.uniform(0, 4.0, N * N)
    b = problems.plane_wave(N, problems.KH, 0.3).reshape(-1)
    xr, info = op.gmres_refine(chi, b, inner_restart=50, inner_tol=1e-4, outer_tol=1e-13)
    xc, ic = op.gmres(chi, b, restart=300, tol=1e-14, max_cycles=3)
    assert info["converged"] and rel(xr, xc) < 1e-10


def test_c64_solver_close_to_c128(orc):
    N = 32
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH))
    chi, b, _ = problems.config_plain(N, seed=1)
    chi = np.minimum(chi, 1.25).reshape(-1)
    x64, i64 = op.gmres(chi, b, restart=30, tol=1e-6, max_cycles=40, precision="c64")
    x128, _ = op.gmres(chi, b, restart=200, tol=1e-12, max_cycles=3)
    assert i64["converged"] and rel(x64, x128) < 2e-5
