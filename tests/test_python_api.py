"""The reference Python module's functions (cugmres_amd.gmres) against the
oracle, in the reference module's orientation."""
import numpy as np
import pytest

from conftest import golden, rel
from cugmres_amd import formats
from cugmres_amd import gmres as pg


def test_plane_wave_orientation():
    E = pg.get_plane_wave(0.3, 5)
    assert E.shape == (5, 5)
    assert np.allclose(E[2, :], np.exp(-1j * 0.3 * np.arange(5)))   # varies along columns
    assert np.allclose(E[:, 3], E[0, 3])


def test_greenfun_matches_scipy_values():
    rows = golden("hankel_scipy.json")["rows"]
    x = np.array([r[0] for r in rows])
    ref = 0.25j * (np.array([r[1] for r in rows]) + 1j * np.array([r[3] for r in rows]))
    got = pg.get_greenfun(x / 2.0, 2.0)
    assert np.max(np.abs(got - ref) / np.maximum(np.abs(ref), 1.0)) < 2e-13


def test_green_matrix_is_python_orientation(orc):
    N, k = 12, 2 * 3.14 / (12 / 6)
    assert rel(pg.get_green_matrix(k, N), orc.shifted_green_block(N, k, 0.0, 0.5)) < 1e-14
    # transposed relative to the C++ orientation
    assert rel(pg.get_green_matrix(k, N).T, orc.shifted_green_block(N, k, 0.5, 0.0)) < 1e-14


def test_toeplitz_embedding_and_eps(orc):
    rng = np.random.default_rng(0)
    N = 7
    blk = rng.standard_normal((N, N)) + 1j * rng.standard_normal((N, N))
    assert np.array_equal(pg.get_toeplitz_mat(blk), orc.embed_block(N, 2 * N - 1, blk))
    mask = rng.integers(0, 2, (N, N))
    assert np.array_equal(pg.get_eps_from_mask(2.25, mask), (1.25 * mask).reshape(-1, 1))
    # the reference's own embedded G_prep has exactly this structure
    g = golden("gprep.npz")["g32"].reshape(63, 63)
    assert np.array_equal(pg.get_toeplitz_mat(g[:32, :32]), g)


def test_complex_array_round_trip(tmp_path):
    a = (np.arange(12) * 0.37 - 1j * np.arange(12) ** 1.5).astype(np.complex64)
    p = tmp_path / "v.txt"
    formats.write_complex(p, a)
    assert np.allclose(pg.get_complex_array(str(p)), a, rtol=1e-7, atol=1e-7)


@pytest.mark.gpu
def test_python_matvec_parity(orc):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    N = 96
    k = 2 * 3.14 / (N / 6)
    x, y = np.mgrid[0:N, 0:N]
    mask = np.zeros((N, N))
    mask[(y - N / 3) ** 2 + (x - N / 2) ** 2 <= (N / 6) ** 2] = 1
    eps = pg.get_eps_from_mask(2.25, mask)
    v = pg.get_plane_wave(k, N).reshape(-1)
    op = orc.Op(N, pg.get_green_matrix(k, N))
    chi = (k ** 2 * eps).reshape(-1)
    assert rel(pg.matvec(v, eps, k), op.matvec(chi, v, 0)) < 2e-6
    assert rel(pg.matvec(v, eps, k, transpose=True), op.matvec(chi, v, 1)) < 2e-6
    assert rel(pg.G_matvec(v, k), op.matvec(chi, v, 2)) < 2e-6
    # g given explicitly (float32 embedding as the reference stores it)
    G = pg.get_toeplitz_mat(pg.get_green_matrix(k, N))
    assert rel(pg.matvec(v, eps, k, g=G), op.matvec(chi, v, 0)) < 2e-6
    blk, G2 = pg.G_matvec(v, k, qweq=True)
    assert blk.shape == (N, N) and G2.shape == (2 * N - 1, 2 * N - 1)


@pytest.mark.gpu
def test_python_gmres_reference_problem(orc):
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    N = 256
    k = 2 * 3.14 / (N / 6)
    x, y = np.mgrid[0:N, 0:N]
    mask = np.zeros((N, N))
    mask[(y - N / 3) ** 2 + (x - N / 2) ** 2 <= (N / 6) ** 2] = 1
    eps = pg.get_eps_from_mask(2.25, mask)
    x0 = pg.get_plane_wave(k, N).reshape(-1)
    xs, info = pg.gmres(x0, eps, k, restart=50, tol=1e-6, max_cycles=10)
    assert info["converged"]
    op = orc.Op(N, pg.get_green_matrix(k, N))
    xc, _ = op.gmres((k ** 2 * eps).reshape(-1), x0, x0=x0, restart=100, tol=1e-12, max_cycles=3)
    assert rel(xs.reshape(-1), xc) < 1e-5
    # the reference's relative error vs the Mie field at N = 256 (Python
    # orientation = transpose of the C++ file), final value of
    # Results/relative_errors_256_0.npy
    ana = golden("mkl_solution.npz")["analytical"].reshape(N, N).T.reshape(-1)
    err = golden("python_histories.npz")["relative_errors_256"][-1]
    assert abs(rel(xs.reshape(-1), ana) - err) < 5e-4
