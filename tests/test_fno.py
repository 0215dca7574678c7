"""The learned preconditioner: the float64 CPU oracle (oracle/fno_ref.py) pinned
against explicit DFT sums, a closed form and finite differences, then the
library's BF16 MFMA forward and backward kernels and cgm_fgmres against it.

BF16 tolerances: activations carry 8 significant bits; the network's
correction z - v is compared at relative L2 1.5e-2 and every gradient block at
2e-2 (measured: about 5e-3 for both); a transposed or dropped term moves these
errors to O(1).
"""
import math

import numpy as np
import pytest
import torch

from conftest import rel
from oracle import fno_ref


def _rand_h(N, B=1, seed=0):
    g = torch.Generator().manual_seed(seed)
    return torch.randn(B, N, N, fno_ref.C, generator=g, dtype=torch.float64)


def _rand_S(seed=1):
    g = torch.Generator().manual_seed(seed)
    s = torch.randn(fno_ref.KX, fno_ref.KY, fno_ref.C, fno_ref.C, 2, generator=g, dtype=torch.float64)
    return torch.complex(s[..., 0], s[..., 1])


def _problem(N, B, seed):
    from cugmres_amd import fno, problems
    rng = np.random.default_rng(seed)
    chi = np.stack([problems.block_eps_map(N, 4, 0.5, 0.75, seed=seed + b) - 1.0 for b in range(B)])
    v = rng.normal(size=(B, N, N)) + 1j * rng.normal(size=(B, N, N))
    v /= np.linalg.norm(v.reshape(B, -1), axis=1)[:, None, None]
    params = fno.init_params(seed, zero_out=False)
    return chi, v, params


# ---- CPU: the oracle -----------------------------------------------------------------
def test_layout_matches_library():
    from cugmres_amd import fno
    from cugmres_amd._lib import lib
    import ctypes
    n = ctypes.c_size_t()
    assert lib().cgm_fno_num_params(ctypes.byref(n)) == 0
    assert n.value == fno.NUM_PARAMS == fno_ref.NUM_PARAMS
    lay = fno.layout()
    end = max(off + int(np.prod(shape)) for off, shape in lay.values())
    assert end == fno.NUM_PARAMS
    assert sum(int(np.prod(shape)) for _, shape in lay.values()) == fno.NUM_PARAMS


def test_oracle_spectral_conv_equals_dense_dft():
    N = 32
    h = _rand_h(N, B=2)
    S = _rand_S()
    a = fno_ref.spectral_conv(h, S)
    b = fno_ref.spectral_conv_dense(h, S)
    assert rel(a.numpy(), b.numpy()) < 1e-12


def test_oracle_single_mode_closed_form():
    # h = cos(2 pi (kx x + ky y)/N) in channel c0 and one mixing weight w
    # (c0 -> o0) at that mode: K(h)[., ., o0] = |w| cos(theta + arg w), all
    # other channels zero.  Pins orientation (in -> out), sign and scale.
    N, kx, ky, c0, o0 = 64, 3, 5, 7, 11
    x = torch.arange(N, dtype=torch.float64)
    th = 2 * math.pi * (kx * x[:, None] + ky * x[None, :]) / N
    h = torch.zeros(1, N, N, fno_ref.C, dtype=torch.float64)
    h[0, :, :, c0] = torch.cos(th)
    S = torch.zeros(fno_ref.KX, fno_ref.KY, fno_ref.C, fno_ref.C, dtype=torch.complex128)
    w = 0.7 - 0.4j
    S[kx, ky, c0, o0] = w
    out = fno_ref.spectral_conv(h, S)
    expect = abs(w) * torch.cos(th + math.atan2(w.imag, w.real))
    assert torch.max(torch.abs(out[0, :, :, o0] - expect)) < 1e-12
    out[0, :, :, o0] = 0
    assert torch.max(torch.abs(out)) < 1e-12
    # the same mode read back through the negative-frequency slot kx -> N - kx
    S2 = torch.zeros_like(S)
    S2[fno_ref.KX - kx, ky, c0, o0] = w   # index of kx' = N - kx in {N-16..N-1}
    h2 = torch.zeros_like(h)
    th2 = 2 * math.pi * (-kx * x[:, None] + ky * x[None, :]) / N
    h2[0, :, :, c0] = torch.cos(th2)
    out2 = fno_ref.spectral_conv(h2, S2)
    expect2 = abs(w) * torch.cos(th2 + math.atan2(w.imag, w.real))
    assert torch.max(torch.abs(out2[0, :, :, o0] - expect2)) < 1e-12


def test_oracle_identity_when_projection_is_zero():
    from cugmres_amd import fno
    N = 32
    p = fno.init_params(3, zero_out=True)
    rng = np.random.default_rng(0)
    chi = rng.uniform(0, 11.1, (1, N, N))
    v = rng.normal(size=(1, N, N)) + 1j * rng.normal(size=(1, N, N))
    z = fno_ref.forward_np(p, chi, v, fno.CHI_SCALE)
    assert np.array_equal(z, v)


def test_oracle_gradient_matches_finite_differences():
    from cugmres_amd import fno
    N = 32
    chi, v, p = _problem(N, 1, 5)
    p = p.astype(np.float64)
    rng = np.random.default_rng(7)
    dz = rng.normal(size=v.shape) + 1j * rng.normal(size=v.shape)
    g = fno_ref.grad_np(p, chi, v, fno.CHI_SCALE, dz)

    def loss(q):
        z = fno_ref.forward_np(q, chi, v, fno.CHI_SCALE)
        return float(np.sum(z.real * dz.real + z.imag * dz.imag))

    lay = fno.layout()
    picks = []
    for name in ["W_in", "b_in", "W0", "b1", "S2", "S3", "W3", "W_out", "b_out"]:
        off, shape = lay[name]
        picks += list(off + rng.integers(0, int(np.prod(shape)), 2))
    for k in picks:
        e = 1e-5 * max(1.0, abs(p[k]))
        qp, qm = p.copy(), p.copy()
        qp[k] += e
        qm[k] -= e
        fd = (loss(qp) - loss(qm)) / (2 * e)
        assert abs(fd - g[k]) <= 1e-6 * max(1.0, abs(g[k])), (k, fd, g[k])


# ---- GPU: the library's kernels --------------------------------------------------------
@pytest.fixture(scope="module")
def gpu():
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    from cugmres_amd import fno, solver
    return solver, fno


@pytest.mark.gpu
@pytest.mark.parametrize("N,B", [(64, 2), (128, 1), (256, 2), (128, 4), (256, 8), (256, 16), (512, 16),
                                 (1024, 1)])
def test_fno_forward_matches_oracle(gpu, N, B):
    # B = 1 forms B_x inside the row kernels, B >= 2 with k_invx; N = 1024
    # is the largest grid, whose row kernels hold 146 KiB of LDS; from 2048
    # rows (256, 8) the row kernels take two rows per work-group, from 4096
    # (256, 16) four, from 8192 (512, 16) the cell GEMM of layers 1-2 eight
    solver, fno = gpu
    chi, v, p = _problem(N, B, 11)
    ctx = solver.Context(N)
    net = fno.Preconditioner(ctx, batch=B, params=p)
    chi_d = torch.from_numpy(chi.reshape(B, -1).astype(np.float32)).cuda()
    v_d = torch.from_numpy(v.reshape(B, -1).astype(np.complex64)).cuda()
    z = net.forward(chi_d, v_d).cpu().numpy().reshape(B, N, N)
    zr = fno_ref.forward_np(p, chi, v, fno.CHI_SCALE)
    err = rel(z - v, zr - v)
    print(f"N={N} B={B} forward rel err of the correction {err:.3e}")
    assert err < 1.5e-2


@pytest.mark.gpu
@pytest.mark.parametrize("N,B", [(64, 3), (256, 3), (256, 8), (256, 16), (512, 16)])
def test_fno_forward_batch_equals_single_vectors(gpu, N, B):
    # one vector (inverse x-DFT inside the row kernels) and a batch (k_invx;
    # at 2048 rows two rows per work-group in the cell GEMM) form the same
    # sums in the same order: equal bit for bit
    solver, fno = gpu
    chi, v, p = _problem(N, B, 5)
    ctx = solver.Context(N)
    chi_d = torch.from_numpy(chi.reshape(B, -1).astype(np.float32)).cuda()
    v_d = torch.from_numpy(v.reshape(B, -1).astype(np.complex64)).cuda()
    zb = fno.Preconditioner(ctx, batch=B, params=p).forward(chi_d, v_d)
    net1 = fno.Preconditioner(ctx, batch=1, params=p)
    for k in range(B):
        z1 = net1.forward(chi_d[k:k + 1].contiguous(), v_d[k:k + 1].contiguous())
        assert torch.equal(zb[k:k + 1], z1), k


@pytest.mark.gpu
@pytest.mark.parametrize("N,B", [(64, 2), (128, 1), (128, 4), (256, 8), (256, 16), (512, 16)])
def test_fno_backward_matches_oracle(gpu, N, B):
    """(512, 16) runs 8,192 grid rows, where the training forward pass takes the
    eight-row cell GEMM that writes the pre-activations the backward reads."""
    solver, fno = gpu
    chi, v, p = _problem(N, B, 21)
    rng = np.random.default_rng(3)
    dz = (rng.normal(size=v.shape) + 1j * rng.normal(size=v.shape)) / N
    ctx = solver.Context(N)
    net = fno.Preconditioner(ctx, batch=B, params=p)
    chi_d = torch.from_numpy(chi.reshape(B, -1).astype(np.float32)).cuda()
    v_d = torch.from_numpy(v.reshape(B, -1).astype(np.complex64)).cuda()
    net.forward(chi_d, v_d, train=True)
    dz_d = torch.from_numpy(dz.reshape(B, -1).astype(np.complex64)).cuda()
    g = net.backward(dz_d).cpu().numpy()
    gr = fno_ref.grad_np(p, chi, v, fno.CHI_SCALE, dz)
    for name, (off, shape) in fno.layout().items():
        size = int(np.prod(shape))
        e = rel(g[off:off + size], gr[off:off + size])
        print(f"N={N} {name:6s} grad rel err {e:.3e}")
        assert e < 2e-2, name


def _small_problem(N, seed=0):
    from cugmres_amd import problems
    chi, b, _ = problems.config_plain(N, seed)
    return chi, b


@pytest.mark.gpu
def test_fgmres_identity_matches_gmres(gpu):
    solver, fno = gpu
    from cugmres_amd import problems
    N = 64
    chi, b = _small_problem(N, 2)
    ctx = solver.Context(N)
    g = ctx.spectrum_helmholtz(problems.KH)
    chi_d = ctx.to_device(chi, "f32")
    b_d = ctx.to_device(b, "c64")
    ctx.set_arnoldi_path(1)
    _, i1 = ctx.gmres(g, chi_d, b_d, restart=12, tol=1e-12, max_cycles=3, history=True)
    ctx.set_arnoldi_path(0)
    _, i2 = ctx.fgmres(g, chi_d, b_d, precond=None, restart=12, tol=1e-12, max_cycles=3, stall_ratio=0.0,
                       history=True)
    assert i1["iterations"] == i2["iterations"] == 36
    assert i2["precond_cycles"] == 3 and not i2["fell_back"]
    h1, h2 = i1["history"], i2["history"]
    assert np.max(np.abs(h1 - h2)) <= 1e-5 * h1[0]


@pytest.mark.gpu
def test_fgmres_with_network_reaches_true_tolerance(gpu, orc):
    # moderate contrast (eps 2.25 blocks), where GMRES(20) converges in a few
    # cycles: a random network must not stop FGMRES from reaching the same
    # true residual, and the stall rule must hand over to plain GMRES
    solver, fno = gpu
    from cugmres_amd import problems
    N = 64
    chi = problems.block_eps_map(N, 8, 0.5, 0.75, seed=4, eps_hi=2.25) - 1.0
    b = problems.plane_wave(N, problems.KH, 0.3)
    ctx = solver.Context(N)
    g = ctx.spectrum_helmholtz(problems.KH)
    chi_d = ctx.to_device(chi, "f32")
    b_d = ctx.to_device(b, "c64")
    p = fno.init_params(9, zero_out=False)
    fno.unpack(p)["W_out"][...] *= 0.05
    net = fno.Preconditioner(ctx, params=p)
    x, info = ctx.fgmres(g, chi_d, b_d, precond=net, restart=20, tol=1e-6, max_cycles=100, stall_ratio=0.0)
    print(info)
    assert info["converged"] and info["precond_cycles"] == info["cycles"]
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH), ctx.M)
    chi64 = chi.reshape(-1)
    r = b.reshape(-1) - op.matvec(chi64, x.cpu().numpy().astype(np.complex128))
    assert np.linalg.norm(r) / np.linalg.norm(b) < 2e-6
    x2, info2 = ctx.fgmres(g, chi_d, b_d, precond=net, restart=20, tol=1e-6, max_cycles=100,
                           stall_ratio=1e-12)
    print(info2)
    assert info2["converged"] and info2["fell_back"] and info2["precond_cycles"] == 1
    r2 = b.reshape(-1) - op.matvec(chi64, x2.cpu().numpy().astype(np.complex128))
    assert np.linalg.norm(r2) / np.linalg.norm(b) < 2e-6


@pytest.mark.gpu
@pytest.mark.parametrize("N,least_gain", [(128, 2.0), (256, 2.0)])
def test_trained_network_gains_over_plain_gmres(gpu, orc, N, least_gain):
    """What training buys: a network trained on the GPU by the library's own
    kernels (at the test's N, 2,400 seeded samples) against plain GMRES(30) on
    held-out seeds of the plain workload's kind, at equal iteration counts.
    The untrained network (s_id = 0, W_out = 0) is P = 0, so the baseline
    without it is plain GMRES."""
    solver, fno = gpu
    from cugmres_amd import learn, problems
    print(f"trained network gain N={N}", flush=True)
    ctx = solver.Context(N)
    net, hist = learn.train(ctx, samples=2400, batch=8, lr=1e-3, seed=3)
    first = np.mean([h[1] for h in hist[:10]])
    last = np.mean([h[1] for h in hist[-30:]])
    print(f"training loss {first:.3f} -> {last:.3f}")
    # (the loss is not a clean progress measure: later steps train on vectors
    # further down the residual chain, which are harder)
    g = ctx.spectrum_helmholtz(problems.KH)
    gains = []
    for s in (9001, 9002):
        chi, b, _ = problems.config_plain(N, s)
        chi_d, b_d = ctx.to_device(chi, "f32"), ctx.to_device(b, "c64")
        _, ip = ctx.gmres(g, chi_d, b_d, restart=30, tol=1e-12, max_cycles=4)
        x, il = ctx.fgmres(g, chi_d, b_d, precond=net, restart=30, tol=1e-12, max_cycles=4, stall_ratio=0.0)
        assert ip["iterations"] == il["iterations"] == 120
        gains.append(ip["rel_residual"] / il["rel_residual"])
        print(f"seed {s}: plain {ip['rel_residual']:.3e}  learned {il['rel_residual']:.3e}")
        # the reported residual is the true one
        op = orc.Op(N, orc.baseline_green_block(N, problems.KH), ctx.M)
        r = b.reshape(-1) - op.matvec(chi.reshape(-1), x.cpu().numpy().astype(np.complex128))
        assert abs(np.linalg.norm(r) / np.linalg.norm(b) - il["rel_residual"]) < 1e-3 * il["rel_residual"] + 1e-6
    assert min(gains) > least_gain, gains


@pytest.mark.gpu
def test_validation_based_stop_keeps_the_best_held_out_parameters(gpu):
    """Two-level training (16 x 16 tiles, N = 64) with a held-out set evaluated
    every 20 steps: the returned network's held-out loss is the lowest of the
    curve, it belongs to the recorded step, and training lowered it below the
    untrained network's."""
    solver, fno = gpu
    from cugmres_amd import learn, problems
    N = 64
    ctx = solver.Context(N)
    net, hist = learn.train(ctx, samples=800, batch=8, lr=1e-3, seed=5, tiles_T=16, map_pool=2, val_every=20)
    curve = net.validation
    assert len(curve) == 5, curve
    g = ctx.spectrum_helmholtz(problems.KH)
    chi_v, v_v, tl = learn.validation_set(ctx, g, 8, 16, 5, 3)
    kept = learn.validation_loss(ctx, g, net, chi_v, v_v)
    best_step, best_loss = min(curve, key=lambda c: c[1])
    assert net.best_step == best_step
    assert kept == pytest.approx(best_loss, rel=1e-5)
    untrained = fno.Preconditioner(ctx, batch=8, params=fno.init_params(5, zero_out=True))
    assert kept < 0.9 * learn.validation_loss(ctx, g, untrained, chi_v, v_v)
    tl.close()

