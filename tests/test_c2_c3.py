"""The mixed-precision workload (smooth permittivity, FP64 refinement) and the
batched multi-right-hand-side workload at sizes the FP64 oracle covers.

Neither has a counterpart in the reference, which solves one system per call
in complex64 (B/CUDA_GMRES/GMRES.cuh:22-446); they wrap that solve."""
import math

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def test_c2_refinement_reaches_fp64_accuracy(orc):
    """Smooth map (white noise filtered at sigma = 4 cells, eps in [1, 12.1]),
    N = 64: FP64 refinement whose corrections come from complex64 flexible
    GMRES with the 16 x 16 tile inverses reaches an FP64 true residual of
    1e-12, and the solution is within 1e-10 of the oracle's FP64 GMRES."""
    import torch
    from cugmres_amd import problems, solver, tiles
    N = 64
    print("c2 refinement N=64", flush=True)
    ctx = solver.Context(N, 0)
    g32, g64 = ctx.spectrum_helmholtz(problems.KH, fp64=True)
    chi, b, _ = problems.config_mixed(N, seed=3)
    chi, b = chi.reshape(-1), b.reshape(-1)
    tl = tiles.Tiles(ctx, 16, 1).factor(ctx.block_helmholtz(problems.KH), ctx.to_device(chi, "f32"))
    x, info = ctx.gmres_refine_pc(g32, g64, ctx.to_device(chi, "f64"), ctx.to_device(b, "c128"), tl,
                                  inner_restart=1000, inner_tol=1e-4, inner_max_cycles=1, outer_tol=1e-12,
                                  max_outer=12)
    assert info["converged"], info
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH), ctx.M)
    xs = x.cpu().numpy()
    assert np.linalg.norm(b - op.matvec(chi, xs)) / np.linalg.norm(b) < 2e-12
    xr, ir = op.gmres(chi, b, restart=1500, tol=1e-14, max_cycles=2)
    assert ir["rel_residual"] < 1e-13
    assert rel(xs, xr) < 1e-10
    tl.close()
    ctx.close()


@pytest.mark.parametrize("N", [64, 96, 128])
def test_c3_sixteen_plane_waves_on_one_map(orc, N):
    """One instance of the batched workload: one block map (blocks of 4-8
    cells), 16 plane waves at angles 2 pi k / 16, solved in lockstep with the
    map's tile inverses factored once.  Every right-hand side converges, its
    FP64 residual (oracle matvec) is below 1.5e-6, and two of them equal
    their single solves."""
    import torch
    from cugmres_amd import problems, solver, tiles
    print(f"c3 instance N={N}", flush=True)
    ctx = solver.Context(N, 0)
    g, g64 = ctx.spectrum_helmholtz(problems.KH, fp64=True)
    ctx.set_flexible_fp64(g64)
    blk = ctx.block_helmholtz(problems.KH)
    eps = problems.block_eps_map(N, 4 + N % 5, 0.5, 0.75, seed=N)
    chi = (eps - 1.0).reshape(-1)
    dchi = ctx.to_device(chi, "f32")
    bs = np.stack([problems.plane_wave(N, problems.KH, 2 * math.pi * k / 16).reshape(-1) for k in range(16)])
    db = torch.as_tensor(bs.astype(np.complex64)).to(ctx.device)
    tl = tiles.Tiles(ctx, 16, 1).factor(blk, dchi)
    xb, infos = ctx.fgmres_batched(g, dchi, db, precond=tl, restart=1500, tol=1e-6, max_cycles=2)
    assert all(i["converged"] for i in infos), [i["rel_residual"] for i in infos]
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH), ctx.M)
    X = xb.cpu().numpy().astype(np.complex128)
    for k in range(16):
        r = np.linalg.norm(bs[k] - op.matvec(chi, X[k])) / np.linalg.norm(bs[k])
        assert r < 1.5e-6, (k, r)
    for k in (0, 11):
        x1, i1 = ctx.fgmres(g, dchi, db[k].contiguous(), precond=tl, restart=1500, tol=1e-6, max_cycles=2,
                            stall_ratio=0.0)
        assert i1["iterations"] == infos[k]["iterations"]
        assert rel(x1.cpu().numpy(), X[k]) < 1e-6
    tl.close()
    ctx.close()
