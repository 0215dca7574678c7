"""Batched lockstep GMRES (cgm_gmres_batched) against single solves.

With one work-group shape for both (cgm_set_arnoldi_path 3: 1,024 threads,
or 2: 256 threads) the batched kernels run the same per-system arithmetic as
a single solve (same blocks, same reduction order), so every system must
reproduce its single solve: equal iteration counts, solutions equal to float
rounding (checked at relative L2 1e-6) and histories equal to 1e-6 of ||r0||.
"""
import math

import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def sol():
    import torch
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cugmres_amd import solver
    return solver


def _systems(N, B, eps_hi, seed0, shared=False):
    from cugmres_amd import problems
    chis, bs = [], []
    for k in range(B):
        eps = problems.block_eps_map(N, 8 if N >= 128 else 4, 0.5, 0.75, seed=seed0 + (0 if shared else k),
                                     eps_hi=eps_hi)
        chis.append(eps - 1.0)
        bs.append(problems.plane_wave(N, problems.KH, 2 * math.pi * k / B))
    return np.stack(chis), np.stack(bs)


def _check(sol, N, B, eps_hi, shared, restart, max_cycles, tol=1e-6, path=3):
    import torch
    from cugmres_amd import problems
    ctx = sol.Context(N)
    ctx.set_arnoldi_path(path)
    g = ctx.spectrum_helmholtz(problems.KH)
    chi, b = _systems(N, B, eps_hi, 7, shared)
    chi_d = torch.from_numpy(chi.reshape(B, -1).astype(np.float32)).cuda()
    b_d = torch.from_numpy(b.reshape(B, -1).astype(np.complex64)).cuda()
    chi_arg = chi_d[0].contiguous() if shared else chi_d
    xb, infos = ctx.gmres_batched(g, chi_arg, b_d, restart=restart, tol=tol, max_cycles=max_cycles,
                                  history=True)
    for k in range(B):
        xs, i1 = ctx.gmres(g, chi_d[k].contiguous(), b_d[k].contiguous(), restart=restart, tol=tol,
                           max_cycles=max_cycles, history=True)
        ib = infos[k]
        assert ib["iterations"] == i1["iterations"], (k, ib, i1)
        assert ib["converged"] == i1["converged"]
        assert rel(xb[k].cpu().numpy(), xs.cpu().numpy()) < 1e-6
        h1 = i1["history"]
        # batched history is laid out by cycle slot (c * restart + j)
        hb = ib["history"][: ib["iterations"] + 1] if ib["cycles"] <= 1 else None
        if hb is not None:
            assert np.max(np.abs(hb - h1[: len(hb)])) <= 1e-6 * h1[0]
    return infos


@pytest.mark.parametrize("path", [3, 2])
def test_batched_own_maps_match_single(sol, path):
    infos = _check(sol, 64, 5, 2.25, False, 20, 20, path=path)
    assert all(i["converged"] for i in infos)
    assert len({i["iterations"] for i in infos}) > 1   # systems stop at different steps


def test_batched_multi_rhs_shared_map(sol):
    # 8 plane waves on one map (chi_stride 0); whatever each reaches in 12
    # cycles, it must be what its single solve reaches
    infos = _check(sol, 128, 8, 2.25, True, 30, 12)
    print([(i["iterations"], i["converged"], i["rel_residual"]) for i in infos])


def test_batched_capped_hard_problem(sol):
    # eps = 12.1 blocks: no system converges in 2 cycles; all must still match
    infos = _check(sol, 64, 3, 12.1, False, 16, 2)
    assert not any(i["converged"] for i in infos)


def test_batched_zero_rhs_and_non_pow2(sol):
    import torch
    from cugmres_amd import problems
    N = 64
    ctx = sol.Context(N)
    g = ctx.spectrum_helmholtz(problems.KH)
    chi, b = _systems(N, 3, 2.25, 3)
    b[1] = 0
    chi_d = torch.from_numpy(chi.reshape(3, -1).astype(np.float32)).cuda()
    b_d = torch.from_numpy(b.reshape(3, -1).astype(np.complex64)).cuda()
    x, infos = ctx.gmres_batched(g, chi_d, b_d, restart=20, tol=1e-6, max_cycles=20)
    assert infos[1]["converged"] and infos[1]["iterations"] == 0
    assert float(torch.linalg.vector_norm(x[1])) == 0.0
    assert infos[0]["converged"] and infos[2]["converged"]
    # N = 96 (3 * 2^5) runs the fused kernels; N = 80 has none, and its
    # systems run one after another
    _check(sol, 96, 2, 2.25, False, 20, 20)
    _check(sol, 80, 2, 2.25, False, 20, 20)


def test_batched_graph_replays_match_eager_launches(sol):
    # without history the cycles are captured once and replayed; every replay
    # must do exactly what the one-by-one launches do (same kernels, same
    # order: bitwise equal results), over several cycles with host reads and
    # eager work between the replays
    import torch
    from cugmres_amd import problems
    N, B = 64, 6
    ctx = sol.Context(N)
    g = ctx.spectrum_helmholtz(problems.KH)
    chi, b = _systems(N, B, 2.25, 11)
    chi_d = torch.from_numpy(chi.reshape(B, -1).astype(np.float32)).cuda()
    b_d = torch.from_numpy(b.reshape(B, -1).astype(np.complex64)).cuda()
    ctx.set_batched_graphs(1)
    xg, ig = ctx.gmres_batched(g, chi_d, b_d, restart=10, tol=1e-7, max_cycles=12)
    ctx.set_batched_graphs(0)
    xe, ie = ctx.gmres_batched(g, chi_d, b_d, restart=10, tol=1e-7, max_cycles=12)
    ctx.set_batched_graphs(1)
    assert max(i["cycles"] for i in ig) > 2
    for k in range(B):
        assert ig[k]["iterations"] == ie[k]["iterations"]
        assert ig[k]["converged"] == ie[k]["converged"]
        assert ig[k]["rel_residual"] == ie[k]["rel_residual"]
    assert torch.equal(xg, xe)
    # and a second graph call on the same context (new capture) agrees too
    xg2, ig2 = ctx.gmres_batched(g, chi_d, b_d, restart=10, tol=1e-7, max_cycles=12)
    assert torch.equal(xg2, xe)
    for k in range(B):
        xs, i1 = ctx.gmres(g, chi_d[k].contiguous(), b_d[k].contiguous(), restart=10, tol=1e-7, max_cycles=12)
        assert i1["iterations"] == ig[k]["iterations"]
        assert rel(xg[k].cpu().numpy(), xs.cpu().numpy()) < 1e-6


def test_batched_matvec_equals_single(sol):
    import torch
    from cugmres_amd import problems
    for N in (64, 96, 256):
        ctx = sol.Context(N)
        g = ctx.spectrum_helmholtz(problems.KH)
        chi, b = _systems(N, 4, 12.1, 5)
        chi_d = torch.from_numpy(chi.reshape(4, -1).astype(np.float32)).cuda()
        x_d = torch.from_numpy(b.reshape(4, -1).astype(np.complex64)).cuda()
        for mode in (0, 1, 2):
            yb = ctx.matvec_batched(g, chi_d, x_d, mode)
            ys = ctx.matvec_batched(g, chi_d[0].contiguous(), x_d, mode)   # one shared map
            for k in range(4):
                y1 = ctx.matvec(g, chi_d[k].contiguous(), x_d[k].contiguous(), mode)
                assert torch.equal(yb[k], y1), (N, mode, k)
                y0 = ctx.matvec(g, chi_d[0].contiguous(), x_d[k].contiguous(), mode)
                assert torch.equal(ys[k], y0), (N, mode, k)


@pytest.mark.parametrize("N,B,shared,with_net,path", [(64, 3, False, True, 3), (128, 4, True, True, 2),
                                                     (64, 2, False, False, 3), (512, 4, False, True, 2)])
def test_fgmres_batched_equals_single_solves(sol, N, B, shared, with_net, path):
    """Lockstep flexible GMRES: every system reproduces its own cgm_fgmres
    solve (same network on each vector, same per-system arithmetic)."""
    import torch
    from cugmres_amd import fno, problems
    chis, bs = _systems(N, B, 2.25, 40, shared=shared)
    ctx = sol.Context(N)
    ctx.set_arnoldi_path(path)
    g = ctx.spectrum_helmholtz(problems.KH)
    chi_d = ctx.to_device(chis[0], "f32") if shared else ctx.to_device(chis, "f32").reshape(B, N * N)
    b_d = ctx.to_device(bs, "c64").reshape(B, N * N)
    net = None
    if with_net:
        p = fno.init_params(5, zero_out=False)
        fno.unpack(p)["W_out"][...] *= 0.05
        net = fno.Preconditioner(ctx, batch=B, params=p)
    xb, ib = ctx.fgmres_batched(g, chi_d, b_d, precond=net, restart=12, tol=1e-7, max_cycles=4, history=True)
    for k in range(B):
        chi_k = chi_d if shared else chi_d[k].contiguous()
        xs, isg = ctx.fgmres(g, chi_k, b_d[k].contiguous(), precond=net, restart=12, tol=1e-7, max_cycles=4,
                             stall_ratio=0.0, history=True)
        assert ib[k]["iterations"] == isg["iterations"]
        assert ib[k]["converged"] == isg["converged"]
        assert rel(xb[k].cpu().numpy(), xs.cpu().numpy()) < 1e-5
        h1, h2 = ib[k]["history"][: isg["iterations"] + 1], isg["history"]
        assert np.max(np.abs(h1 - h2)) <= 1e-5 * h2[0]
        assert ib[k]["rel_residual"] == pytest.approx(isg["rel_residual"], rel=1e-4)
    torch.cuda.synchronize()


def test_smaller_batch_after_larger_leaves_memory_past_x_alone(sol):
    """A batched call after a larger one on the same context touches only its
    own systems: rows of the buffer past the caller's x keep their sentinel
    values (the launches are sized by the call's batch, not the workspace's)."""
    import torch
    N = 64
    ctx = sol.Context(N, 0)
    g = ctx.spectrum_helmholtz(__import__("cugmres_amd.problems", fromlist=["KH"]).KH)
    chis, bs = _systems(N, 6, 2.25, 300)
    dchi = torch.as_tensor(chis.reshape(6, -1).astype(np.float32)).to(ctx.device)
    db = torch.as_tensor(bs.reshape(6, -1).astype(np.complex64)).to(ctx.device)
    for fn in ("gmres_batched", "fgmres_batched"):
        getattr(ctx, fn)(g, dchi, db, restart=20, tol=1e-6, max_cycles=3)   # workspace for 6 systems
        buf = torch.full((6, N * N), 7.0 + 3.0j, dtype=torch.complex64, device=ctx.device)
        x = buf[:2]
        getattr(ctx, fn)(g, dchi[:2], db[:2].contiguous(), x.zero_(), restart=20, tol=1e-6, max_cycles=3)
        torch.cuda.synchronize()
        tail = buf[2:].cpu().numpy()
        assert np.all(tail == np.complex64(7.0 + 3.0j)), fn
        assert np.all(np.isfinite(x.cpu().numpy()))
    ctx.close()
