"""Block GMRES over the right-hand sides of one map (cugmres_amd/block.py).

The reference solves one system per call (B/CUDA_GMRES/GMRES.cuh:22-35, one
dot at a time at :289-303); block GMRES has no reference counterpart.  The CPU
tests pin the core on dense systems by what defines it: the solution of a
well-conditioned system equals a direct solve, and after k block steps every
column's residual estimate is the minimum of || b_c - A P z || over z in the
block Krylov space K_k(A P, R0), computed here independently by a QR of the
explicit block Krylov matrix and a least-squares solve.  The GPU test runs the
batched workload's shape (one map, 16 plane waves) through the device path and
checks every solution with the oracle's FP64 matvec and against single
solves."""
import math

import numpy as np
import pytest

from conftest import rel


def _dense(n, seed, scale=0.3):
    import torch
    g = torch.Generator().manual_seed(seed)
    re = torch.randn(n, n, generator=g, dtype=torch.float64)
    im = torch.randn(n, n, generator=g, dtype=torch.float64)
    return torch.eye(n, dtype=torch.complex128) + scale * torch.complex(re, im) / math.sqrt(n)


def _rhs(n, s, seed):
    import torch
    g = torch.Generator().manual_seed(seed)
    return torch.complex(torch.randn(s, n, generator=g, dtype=torch.float64),
                         torch.randn(s, n, generator=g, dtype=torch.float64))


def test_block_gmres_solves_a_dense_system():
    """n = 96, s = 4, no preconditioner, one cycle long enough to span the
    space: the solution equals torch.linalg.solve, within ceil(n / s) block
    steps (exact arithmetic needs at most that many)."""
    import torch
    from cugmres_amd.block import block_gmres
    n, s = 96, 4
    A = _dense(n, 1)
    B = _rhs(n, s, 2)
    X, info = block_gmres(torch, lambda Y: Y @ A.T, lambda Y: Y, B, restart=30, tol=1e-12, max_cycles=2)
    ref = torch.linalg.solve(A, B.T).T
    assert all(info["converged"]), info
    assert rel(X.numpy(), ref.numpy()) < 1e-9
    assert info["block_steps"] <= n // s + 1


@pytest.mark.parametrize("s,precond", [(1, False), (3, False), (3, True)])
def test_block_residual_estimates_are_the_block_krylov_minimum(s, precond):
    """After k block steps the estimate for column c equals
    min || r_c - A P K y || / || b_c ||, K an orthonormal basis of
    [R0, (A P) R0, ..., (A P)^(k-1) R0] built here from the explicit powers."""
    import torch
    from cugmres_amd.block import block_gmres
    n = 60
    A = _dense(n, 3, scale=0.8)
    P = torch.diag(1.0 / torch.diagonal(A)) + 0.05 * _dense(n, 4, scale=0.5) if precond else torch.eye(
        n, dtype=torch.complex128)
    B = _rhs(n, s, 5)
    est = []
    block_gmres(torch, lambda Y: Y @ A.T, lambda Y: Y @ P.T, B, restart=6, tol=1e-30, max_cycles=1,
                on_step=lambda k, e: est.append(np.array(e)))
    AP = (A @ P).numpy()
    R0 = B.numpy().T                                  # x0 = 0: R0 = B, columns
    bn = np.linalg.norm(R0, axis=0)
    blocks = [R0]
    for k in range(1, len(est) + 1):
        K, _ = np.linalg.qr(np.hstack(blocks))
        for c in range(s):
            y, *_ = np.linalg.lstsq(AP @ K, R0[:, c], rcond=None)
            want = np.linalg.norm(R0[:, c] - AP @ K @ y) / bn[c]
            assert abs(est[k - 1][c] - want) <= 1e-9 * max(1.0, want), (k, c, est[k - 1][c], want)
        blocks.append(AP @ blocks[-1])
    assert len(est) == 6


def test_block_gmres_restarts_and_dependent_right_hand_sides():
    """Short cycles (restart 3) and a right-hand side repeated twice (the
    first block is rank-deficient): every column still reaches the direct
    solution."""
    import torch
    from cugmres_amd.block import block_gmres
    n = 80
    A = _dense(n, 6, scale=0.25)
    B = _rhs(n, 3, 7)
    B = torch.cat([B, B[1:2]])
    X, info = block_gmres(torch, lambda Y: Y @ A.T, lambda Y: Y, B, restart=3, tol=1e-11, max_cycles=60)
    ref = torch.linalg.solve(A, B.T).T
    assert all(info["converged"]), info
    assert rel(X.numpy(), ref.numpy()) < 1e-8


@pytest.mark.gpu
def test_block_gmres_on_one_map_matches_single_solves(orc):
    """N = 64, one map of 4-cell blocks (eps in {1, 12.1}), 16 plane waves at
    2 pi k / 16, 16 x 16 tile inverses: block GMRES (FP64 true residual for the
    test and the restarts) converges every right-hand side; the oracle's FP64
    matvec confirms each residual, and the solutions agree with single
    flexible GMRES solves with the same tiles and with the oracle's converged
    solve of two of them."""
    import torch
    from cugmres_amd import block, problems, solver, tiles as gt
    from oracle import tiles
    from test_tiles import _oracle_right_pc_gmres
    N, T, s = 64, 16, 16
    print(f"block gmres N={N} T={T} s={s}", flush=True)
    ctx = solver.Context(N, 0)
    g, g64 = ctx.spectrum_helmholtz(problems.KH, fp64=True)
    ctx.set_flexible_fp64(g64)
    eps = problems.block_eps_map(N, 4, 0.5, 0.75, seed=5)
    chi = (eps - 1.0).reshape(-1)
    dchi = ctx.to_device(chi, "f32")
    dchi64 = ctx.to_device(chi, "f64")
    B = torch.stack([ctx.to_device(problems.plane_wave(N, problems.KH, 2 * math.pi * k / s).reshape(-1))
                     for k in range(s)])
    tl = gt.Tiles(ctx, T, 1).factor(ctx.block_helmholtz(problems.KH), dchi)
    X, info = block.solve_map(ctx, g, dchi, B, tiles=tl, restart=64, tol=1e-6, max_cycles=6, gamma64=g64,
                              chi64=dchi64)
    print("block", info["block_steps"], info["cycles"], info["breakdowns"], max(info["rel_residual"]), flush=True)
    assert all(info["converged"]), info
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH), ctx.M)
    Xs = X.cpu().numpy().astype(np.complex128)
    Bs = B.cpu().numpy().astype(np.complex128)
    for c in range(s):
        r = Bs[c] - op.matvec(chi, Xs[c])
        assert np.linalg.norm(r) / np.linalg.norm(Bs[c]) < 1.5e-6, c
    single_steps = 0
    for c in range(s):
        x, inf = ctx.fgmres(g, dchi, B[c].contiguous(), precond=tl, restart=400, tol=1e-6, max_cycles=4,
                            stall_ratio=0.0)
        assert inf["converged"], (c, inf)
        single_steps += inf["iterations"]
        assert rel(Xs[c], x.cpu().numpy()) < 2e-5, c
    print("single steps", single_steps, flush=True)
    K = tiles.tile_inverses(orc.baseline_green_block(N, problems.KH), chi, N, T)
    for c in (0, 5):
        y, _ = _oracle_right_pc_gmres(lambda v: op.matvec(chi, v), lambda v: tiles.tile_apply(K, v, N, T), Bs[c],
                                      1e-12)
        assert rel(Xs[c], y) < 2e-5, c
    tl.close()
    ctx.close()


@pytest.mark.parametrize("chunk", [0, 64, 96])
def test_gemm_forms_equal_their_definitions(chunk):
    """The chunked batched projection, the real-GEMM update and the chunked
    Gram matrix equal V^H W, W - (V H)^T and conj(W) W^T formed directly (n a
    multiple of the chunk, and not: then one chunk)."""
    import torch
    from cugmres_amd.block import inner, project, update
    g = torch.Generator().manual_seed(11)
    n, k, s = 384, 9, 4
    V = torch.complex(torch.randn(k, n, generator=g, dtype=torch.float64),
                      torch.randn(k, n, generator=g, dtype=torch.float64))
    W = torch.complex(torch.randn(s, n, generator=g, dtype=torch.float64),
                      torch.randn(s, n, generator=g, dtype=torch.float64))
    H = V.conj() @ W.T
    assert rel(project(torch, V, W, chunk=chunk).resolve_conj().numpy(), H.numpy()) < 1e-13
    assert rel(update(torch, W, V, H).numpy(), (W - (V.T @ H).T).numpy()) < 1e-13
    assert rel(inner(torch, W, W, chunk=chunk).numpy(), (W.conj() @ W.T).numpy()) < 1e-13


def test_solve_map_rejects_wrong_arguments_before_any_device_work():
    """Shapes and dtypes are checked first: a batch of maps, a complex128
    right-hand side block, a wrong length or half of the FP64 pair raise
    ValueError without touching the (here absent) device."""
    import torch
    from cugmres_amd.block import solve_map

    class Ctx:
        N = 8

    Ctx.torch = torch
    n = 64
    B = torch.zeros(3, n, dtype=torch.complex64)
    chi = torch.zeros(n, dtype=torch.float32)
    for args, kw in [((B.to(torch.complex128), chi), {}), ((B[:, :10], chi), {}),
                     ((B, torch.zeros(2, n, dtype=torch.float32)), {}), ((B, chi.double()), {}),
                     ((B, chi), {"gamma64": object()})]:
        with pytest.raises(ValueError):
            solve_map(Ctx(), None, args[1], args[0], **kw)


def test_block_gmres_on_the_helmholtz_operator_with_oracle_tiles(orc):
    """The method on the physics, in FP64 on the CPU: the oracle's operator
    (N = 16, 4-cell block map, eps in {1, 12.1}) and its 8 x 8 tile inverses
    as P, 4 plane waves at 2 pi k / 4; block GMRES reaches 1e-10 and its
    solutions equal a dense solve of the operator assembled column by column
    from the oracle's matvec."""
    import torch
    from cugmres_amd import problems
    from cugmres_amd.block import block_gmres
    from oracle import tiles
    N, T, s = 16, 8, 4
    n = N * N
    eps = problems.block_eps_map(N, 4, 0.5, 0.75, seed=1)
    chi = (eps - 1.0).reshape(-1)
    blk = orc.baseline_green_block(N, problems.KH)
    op = orc.Op(N, blk)
    K = tiles.tile_inverses(blk, chi, N, T)
    B = np.stack([problems.plane_wave(N, problems.KH, 2 * math.pi * k / s).reshape(-1) for k in range(s)])

    def rows(f):
        return lambda Y: torch.as_tensor(np.stack([f(y) for y in Y.numpy()]))

    X, info = block_gmres(torch, rows(lambda y: op.matvec(chi, y)), rows(lambda y: tiles.tile_apply(K, y, N, T)),
                          torch.as_tensor(B.astype(np.complex128)), restart=40, tol=1e-10, max_cycles=3)
    assert all(info["converged"]), info
    A = np.stack([op.matvec(chi, e) for e in np.eye(n, dtype=np.complex128)], axis=1)
    ref = np.linalg.solve(A, B.T).T
    assert rel(X.numpy(), ref) < 1e-8
    assert info["block_steps"] < n // s
