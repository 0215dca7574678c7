"""GPU parity tests: the HIP path through the C ABI against the CPU oracle.

Tolerances: complex64 operator and Krylov vectors (relative L2 2e-6 for one
matvec, 1e-5 for converged solutions against the reference solve), complex128 paths
1e-12 / 1e-10.
"""
import ctypes
import math

import numpy as np
import pytest

from conftest import golden, rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sol():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cugmres_amd import solver
    return solver


_ctx_cache = {}


def ctx_for(sol, N):
    if N not in _ctx_cache:
        _ctx_cache[N] = sol.Context(N)
    return _ctx_cache[N]


def quarter(ghat, M):
    H = M // 2 + 1
    return ghat[:H, :H].T.reshape(-1)  # column-frequency major


def np_(t):
    return t.detach().cpu().numpy()


# ---- spectra -----------------------------------------------------------------------
@pytest.mark.parametrize("N", [6, 16, 48, 100, 256, 512])
def test_spectrum_helmholtz_parity(sol, orc, N):
    from cugmres_amd import problems
    c = ctx_for(sol, N)
    g32, g64 = c.spectrum_helmholtz(problems.KH, fp64=True)
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH), c.M)
    ref = quarter(op.ghat, c.M)
    assert rel(np_(g64), ref) < 1e-12
    assert rel(np_(g32), ref) < 1e-6


def test_spectrum_from_gprep_and_reference(sol, orc):
    N = 64
    c = ctx_for(sol, N)
    gp = golden("gprep.npz")["g64"]
    g = c.spectrum_from_gprep(gp)
    blk, _ = orc.gprep_to_block(N, gp)
    ref = quarter(orc.Op(N, blk, c.M).ghat, c.M)
    assert rel(np_(g), ref) < 1e-6
    # device-generated reference operator (FP64 Bessel, float-free k)
    k = 2 * 3.14 / (N / 6)
    _, g64 = c.spectrum_reference(k, 0.5, 0.0, fp64=True)
    ref2 = quarter(orc.Op(N, orc.shifted_green_block(N, k, 0.5, 0.0), c.M).ghat, c.M)
    assert rel(np_(g64), ref2) < 1e-12


def test_spectrum_rejects_non_symmetric(sol):
    from cugmres_amd import _lib
    N = 16
    c = ctx_for(sol, N)
    gp = np.random.default_rng(0).standard_normal((2 * N - 1) ** 2).astype(np.complex64)
    with pytest.raises(_lib.CgmError) as e:
        c.spectrum_from_gprep(gp)
    assert e.value.status == _lib.ERROR_NOT_SYMMETRIC


# ---- operator -----------------------------------------------------------------------
# one N per supported embedding size M (16 ... 4096)
MATVEC_NS = [6, 12, 16, 20, 32, 48, 64, 96, 128, 192, 256, 384, 512, 600, 1024, 1500, 2048]


@pytest.mark.parametrize("N", MATVEC_NS)
def test_matvec_parity_all_embeddings(sol, orc, N):
    from cugmres_amd import problems
    c = ctx_for(sol, N)
    g32 = c.spectrum_helmholtz(problems.KH)
    rng = np.random.default_rng(N)
    chi = rng.uniform(0, 11.1, N * N)
    x = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH), c.M)
    dchi, dx = c.to_device(chi, "f32"), c.to_device(x, "c64")
    modes = (0, 1, 2) if N <= 512 else (0,)
    for mode in modes:
        y = np_(c.matvec(g32, dchi, dx, mode))
        ref = op.matvec(chi.astype(np.float32).astype(np.float64),
                        x.astype(np.complex64).astype(np.complex128), mode)
        assert rel(y, ref) < 2e-6, (N, mode)


@pytest.mark.parametrize("N", [20, 256])
def test_matvec_c128_parity(sol, orc, N):
    from cugmres_amd import problems
    c = ctx_for(sol, N)
    _, g64 = c.spectrum_helmholtz(problems.KH, fp64=True)
    rng = np.random.default_rng(N + 1)
    chi = rng.uniform(0, 11.1, N * N)
    x = rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N)
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH), c.M)
    for mode in (0, 1):
        y = np_(c.matvec_c128(g64, c.to_device(chi, "f64"), c.to_device(x, "c128"), mode))
        assert rel(y, op.matvec(chi, x, mode)) < 1e-12


def test_adjoint_identity_gpu(sol):
    from cugmres_amd import problems
    N = 128
    c = ctx_for(sol, N)
    g = c.spectrum_helmholtz(problems.KH)
    rng = np.random.default_rng(9)
    chi = c.to_device(rng.uniform(0, 11.1, N * N), "f32")
    x = c.to_device(rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N))
    y = c.to_device(rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N))
    lhs = np.sum(np_(c.matvec(g, chi, x, 0)).astype(np.complex128) * np_(y))
    rhs = np.sum(np_(x).astype(np.complex128) * np_(c.matvec(g, chi, y, 1)))
    assert abs(lhs - rhs) < 1e-5 * abs(lhs)


# ---- GMRES ---------------------------------------------------------------------------
@pytest.fixture(scope="module")
def ref256_gpu(sol, orc):
    N = 256
    c = ctx_for(sol, N)
    gp = orc.ref_gprep(N)
    g = c.spectrum_from_gprep(gp)
    mask = orc.ref_cylinder_mask(N).reshape(-1)
    chi = (np.float32(orc.ref_chi(N)) * mask).astype(np.float32)
    b = orc.ref_plane_wave(N).reshape(-1)
    blk, _ = orc.gprep_to_block(N, gp)
    op = orc.Op(N, blk)
    xc, info = op.gmres(chi.astype(np.float64), b, x0=b, restart=100, tol=1e-12, max_cycles=3)
    assert info["converged"]
    return c, g, chi, b, xc


def test_gmres_reference_problem(sol, ref256_gpu):
    c, g, chi, b, xc = ref256_gpu
    x = c.to_device(b)
    x, info = c.gmres(g, c.to_device(chi, "f32"), c.to_device(b), x, restart=30, tol=1e-6,
                      max_cycles=20, history=True)
    assert info["converged"], info
    xs = np_(x)
    assert rel(xs, xc) < 1e-5
    assert rel(xs, golden("mkl_solution.npz")["solution"]) < 3e-5
    h = info["history"]
    assert h[0] == pytest.approx(info["beta0"], rel=1e-12)


def test_gmres_deterministic(sol, ref256_gpu):
    c, g, chi, b, _ = ref256_gpu
    outs = []
    for _ in range(2):
        x = c.to_device(b)
        x, _ = c.gmres(g, c.to_device(chi, "f32"), c.to_device(b), x, restart=30, tol=1e-6,
                       max_cycles=20)
        outs.append(np_(x))
    assert np.array_equal(outs[0], outs[1])


def test_compat_fast_gmres_matches_mkl_solution(sol, orc, ref256_gpu):
    from cugmres_amd import _lib
    c, g, chi, b, xc = ref256_gpu
    import torch
    N = 256
    mask = torch.as_tensor(orc.ref_cylinder_mask(N).reshape(-1).astype(np.bool_)).to(c.device)
    x = c.to_device(b)
    resid = ctypes.c_void_p()
    n_out = ctypes.c_uint()
    rvt = ctypes.c_bool()
    nts = sol.n_timestamps(44)
    times = (ctypes.c_long * (2 * nts))()
    st = _lib.lib().cgm_Fast_GMRES_with_CUDA(
        ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(mask.data_ptr()), ctypes.c_void_p(x.data_ptr()),
        ctypes.byref(resid), ctypes.byref(n_out), c.handle, ctypes.c_float(0.001), False, 0,
        ctypes.byref(rvt), 44, ctypes.cast(times, ctypes.c_void_p), N)
    assert st == 0
    assert n_out.value == 44
    res = np.zeros(45, np.float32)
    _lib.lib().cgm_memcpy_d2h(res.ctypes.data_as(ctypes.c_void_p), resid, 45 * 4)
    _lib.lib().cgm_free(resid)
    assert res[0] == pytest.approx(518.9833, rel=1e-5)   # ||b - A b||, oracle and MKL dump
    assert np.all(np.diff(res) <= 1e-6 * res[0])          # unitary rotations: monotone
    xs = np_(x)
    assert rel(xs, golden("mkl_solution.npz")["solution"]) < 3e-5
    assert rel(xs, xc) < 2e-5
    assert rvt.value == (res[44] > 0.001)


def test_compat_matvec_and_lu(sol, orc, ref256_gpu):
    from cugmres_amd import _lib
    import torch
    c, g, chi, b, _ = ref256_gpu
    N = 256
    mask = torch.as_tensor(orc.ref_cylinder_mask(N).reshape(-1).astype(np.bool_)).to(c.device)
    x = c.to_device(b)
    y = torch.empty_like(x)
    blk, _ = orc.gprep_to_block(N, orc.ref_gprep(N))
    op = orc.Op(N, blk)
    for fn, mode in (("cgm_G_x_fft_matvec", 0), ("cgm_G_x_fft_matvec_T", 1)):
        st = getattr(_lib.lib(), fn)(ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(mask.data_ptr()),
                                     ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                     c.handle, N)
        assert st == 0
        assert rel(np_(y), op.matvec(chi.astype(np.float64), b.astype(np.complex128), mode)) < 2e-6
    # cuSolve_LES: column-major m x m complex64 system
    m = 12
    rng = np.random.default_rng(4)
    A = (rng.standard_normal((m, m)) + 1j * rng.standard_normal((m, m))).astype(np.complex64)
    B = (rng.standard_normal(m) + 1j * rng.standard_normal(m)).astype(np.complex64)
    dA = torch.as_tensor(A.T.reshape(-1).copy()).to(c.device)
    dB = torch.as_tensor(B.copy()).to(c.device)
    ipiv = torch.zeros(m, dtype=torch.int32, device=c.device)
    info = torch.zeros(1, dtype=torch.int32, device=c.device)
    st = _lib.lib().cgm_cuSolve_LES(ctypes.c_void_p(dA.data_ptr()), ctypes.c_void_p(dB.data_ptr()), m,
                                     c.handle, None, ctypes.c_void_p(ipiv.data_ptr()),
                                     ctypes.c_void_p(info.data_ptr()))
    assert st == 0 and int(info.item()) == 0
    ref = np.linalg.solve(A.astype(np.complex128), B.astype(np.complex128))
    assert rel(np_(dB), ref) < 1e-5


def test_gmres_baseline_operator_small(sol, orc):
    """North-star operator with a contrast that converges (eps = 2.25 blocks)."""
    from cugmres_amd import problems
    N = 64
    c = ctx_for(sol, N)
    g = c.spectrum_helmholtz(problems.KH)
    rng = problems.SplitMix64(3)
    th = 2 * math.pi * rng.uniform()
    eps = problems.block_eps_map(N, 4, 0.5, 1.0, rng=rng, eps_hi=2.25)
    chi = (eps - 1).reshape(-1)
    b = problems.plane_wave(N, problems.KH, th).reshape(-1)
    x, info = c.gmres(g, c.to_device(chi, "f32"), c.to_device(b), restart=30, tol=1e-6,
                      max_cycles=30)
    assert info["converged"]
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH), c.M)
    xc, ic = op.gmres(chi, b, restart=200, tol=1e-12, max_cycles=3)
    assert rel(np_(x), xc) < 1e-5
    # the reported true residual is the oracle's residual of the GPU iterate
    r = b - op.matvec(chi, np_(x).astype(np.complex128))
    assert np.linalg.norm(r) / np.linalg.norm(b) == pytest.approx(info["rel_residual"], rel=2e-2, abs=2e-7)
    # adjoint solve
    xa, ia = c.gmres(g, c.to_device(chi, "f32"), c.to_device(b), restart=30, tol=1e-6,
                     max_cycles=30, adjoint=True)
    xac, _ = op.gmres(chi, b, restart=200, tol=1e-12, max_cycles=3, adjoint=True)
    assert ia["converged"] and rel(np_(xa), xac) < 1e-5


def test_identity_operator_and_zero_rhs(sol):
    from cugmres_amd import problems
    N = 32
    c = ctx_for(sol, N)
    g = c.spectrum_helmholtz(problems.KH)
    b = problems.plane_wave(N).reshape(-1)
    x, info = c.gmres(g, c.to_device(np.zeros(N * N), "f32"), c.to_device(b), restart=10, tol=1e-6)
    assert info["converged"] and info["iterations"] == 1
    assert rel(np_(x), b) < 1e-6
    z, info = c.gmres(g, c.to_device(np.ones(N * N), "f32"), c.to_device(np.zeros(N * N)), restart=10)
    assert info["converged"] and info["iterations"] == 0 and not np.any(np_(z))


def test_nonfinite_input_is_reported(sol):
    from cugmres_amd import _lib, problems
    N = 32
    c = ctx_for(sol, N)
    g = c.spectrum_helmholtz(problems.KH)
    b = problems.plane_wave(N).reshape(-1)
    b[5] = np.nan
    with pytest.raises(_lib.CgmError) as e:
        c.gmres(g, c.to_device(np.ones(N * N), "f32"), c.to_device(b), restart=10)
    assert e.value.status == _lib.ERROR_NONFINITE


def test_refinement_fp64(sol, orc):
    from cugmres_amd import problems
    N = 64
    c = ctx_for(sol, N)
    g32, g64 = c.spectrum_helmholtz(problems.KH, fp64=True)
    rng = np.random.default_rng(12)
    chi = rng.uniform(0, 1.25, N * N)
    b = problems.plane_wave(N, problems.KH, 0.4).reshape(-1)
    x, info = c.gmres_refine(g32, g64, c.to_device(chi, "f64"), c.to_device(b, "c128"),
                             inner_restart=50, inner_tol=1e-4, outer_tol=1e-12)
    assert info["converged"], info
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH), c.M)
    xc, _ = op.gmres(chi, b, restart=300, tol=1e-14, max_cycles=3)
    assert rel(np_(x), xc) < 1e-10


def test_plain_config_full_size_invariants(sol, orc):
    """N = 512 benchmark workload, GMRES(30): true residual is non-increasing
    over restarts and equals the oracle's residual of the returned iterate."""
    from cugmres_amd import problems
    N = 512
    c = ctx_for(sol, N)
    g = c.spectrum_helmholtz(problems.KH)
    chi, b, th = problems.config_plain(N, seed=0)
    chi = chi.reshape(-1)
    b = b.reshape(-1)
    dchi, db = c.to_device(chi, "f32"), c.to_device(b)
    x = None
    last = 1.0
    for _ in range(3):
        x, info = c.gmres(g, dchi, db, x, restart=30, tol=1e-6, max_cycles=2)
        assert info["rel_residual"] <= last * (1 + 1e-6)
        last = info["rel_residual"]
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH), c.M)
    r = b - op.matvec(chi, np_(x).astype(np.complex128))
    assert np.linalg.norm(r) / np.linalg.norm(b) == pytest.approx(last, rel=1e-3)


@pytest.mark.parametrize("N,restart,cycles", [(64, 30, 20), (128, 20, 20), (256, 30, 20), (512, 30, 3),
                                              (1024, 30, 3), (2048, 30, 2), (96, 30, 20), (192, 30, 20),
                                              (384, 30, 3), (768, 30, 2)])
def test_fused_and_classic_arnoldi_agree(sol, orc, N, restart, cycles):
    """Fused lagged-CGS2 steps (1,024-thread work-groups, and 256-thread ones
    where they exist) and the two-pass CGS2 kernels give the same iterates,
    iteration counts and Givens estimates (converged solves at the small
    sizes, a fixed number of cycles at the large ones)."""
    from cugmres_amd import problems
    c = ctx_for(sol, N)
    g = c.spectrum_helmholtz(problems.KH)
    rng = problems.SplitMix64(N)
    th = 2 * math.pi * rng.uniform()
    eps = problems.block_eps_map(N, 4, 0.5, 1.0, rng=rng, eps_hi=1.6)
    chi = c.to_device((eps - 1).reshape(-1), "f32")
    b = c.to_device(problems.plane_wave(N, problems.KH, th).reshape(-1))
    paths = (3, 1, 2) if N in (64, 96, 128, 192, 256, 512) else (3, 1)
    out = []
    for path in paths:
        c.set_arnoldi_path(path)
        x, info = c.gmres(g, chi, b, restart=restart, tol=1e-6, max_cycles=cycles, history=True)
        out.append((np_(x), info))
    c.set_arnoldi_path(0)
    for (x0, i0) in out[:1] + out[2:]:
        _agree(N, restart, cycles, x0, i0, *out[1])


def _agree(N, restart, cycles, x0, i0, x1, i1):
    if N <= 256:
        assert i0["converged"] and i1["converged"]
        assert rel(x0, x1) < 2e-5
    else:
        assert i0["iterations"] == i1["iterations"] == restart * cycles
        assert rel(x0, x1) < 1e-4
        assert i0["rel_residual"] == pytest.approx(i1["rel_residual"], rel=1e-3)
    assert abs(i0["iterations"] - i1["iterations"]) <= 1
    k = min(len(i0["history"]), len(i1["history"]))
    h0 = i0["history"][0]
    assert np.allclose(i0["history"][:k], i1["history"][:k], rtol=1e-3, atol=1e-5 * h0)


def test_autotuned_tiles_are_legal_and_keep_parity(sol, orc):
    """The autotuner picks tiles within the kernels' limits, and a matvec with
    every candidate tile gives the same result."""
    from cugmres_amd import problems
    N = 512
    c = ctx_for(sol, N)
    tiles, us = c.tiles()
    assert all(t in (1, 2, 4, 8) for t in tiles) and all(u > 0 for u in us)
    g = c.spectrum_helmholtz(problems.KH)
    rng = np.random.default_rng(3)
    chi = c.to_device(rng.uniform(0, 11.1, N * N), "f32")
    x = c.to_device(rng.standard_normal(N * N) + 1j * rng.standard_normal(N * N))
    y0 = np_(c.matvec(g, chi, x))
    c.autotune()
    assert rel(np_(c.matvec(g, chi, x)), y0) < 1e-6
