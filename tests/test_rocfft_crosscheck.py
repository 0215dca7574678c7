"""Cross-check of the hand-written FFT matvec against rocFFT.

The library's matvec (kernel spectrum precomputed in FP64 by its own FFT,
three fused LDS Stockham passes in complex64) is compared with the same
circulant-embedding product formed entirely by torch.fft on the device
(hipFFT/rocFFT, complex128): the generator block from cgm_block_helmholtz is
embedded into the M x M circulant, transformed, multiplied and transformed
back, which is what the reference does with cuFFT
(B/CUDA_GMRES/launch_GMRES.cuh:138-142, B/CUDA_GMRES/Fast_matvec.cuh:13-26).
Every embedding size of the supported grids is covered: M = 2N = 128 ... 4096
(powers of two) and M = 96 ... 1536 (3 * 2^k)."""
import numpy as np
import pytest

from conftest import rel

pytestmark = pytest.mark.gpu


def _torch_matvec(torch, blk, chi, x, N, M, mode):
    """x - G(chi x) (mode 0), x - chi G(x) (1) or G x (2) with G applied as the
    M x M circulant embedding of the even block g[|di|, |dj|], all complex128
    through torch.fft."""
    g = blk.reshape(N, N)
    idx = torch.arange(M, device=x.device)
    off = torch.where(idx < M - idx, idx, M - idx)          # |signed offset| on the circle
    valid = off < N
    offc = torch.clamp(off, max=N - 1)
    E = g[offc[:, None], offc[None, :]] * (valid[:, None] & valid[None, :])
    gh = torch.fft.fft2(E)
    u = (chi * x) if mode == 0 else x
    P = torch.zeros((M, M), dtype=torch.complex128, device=x.device)
    P[:N, :N] = u.reshape(N, N)
    y = torch.fft.ifft2(gh * torch.fft.fft2(P))[:N, :N].reshape(-1)
    if mode == 2:
        return y
    return x - (y if mode == 0 else chi * y)


@pytest.mark.parametrize("N", [64, 128, 256, 512, 1024, 2048, 48, 96, 192, 384, 768])
def test_matvec_matches_rocfft(N):
    import torch
    from cugmres_amd import problems, solver
    print(f"rocfft cross-check N={N}", flush=True)
    ctx = solver.Context(N, 0)
    assert ctx.M == (2 * N)
    g = ctx.spectrum_helmholtz(problems.KH)
    blk = ctx.block_helmholtz(problems.KH)
    chi, b, _ = problems.config_plain(N, seed=N)
    x32 = ctx.to_device(b.reshape(-1))
    c32 = ctx.to_device(chi.reshape(-1), "f32")
    x64 = x32.to(torch.complex128)
    c64 = c32.to(torch.float64)
    for mode in (0, 1, 2):
        y = ctx.matvec(g, c32 if mode != 2 else None, x32, mode)
        ref = _torch_matvec(torch, blk, c64, x64, N, ctx.M, mode)
        err = rel(y.cpu().numpy(), ref.cpu().numpy())
        assert err < 2e-6, (N, mode, err)
    ctx.close()


@pytest.mark.parametrize("wide", [0, 1])
def test_batched_matvec_matches_rocfft_with_both_radix_plans(wide):
    """N = 512 (M = 1024): the LDS-staged radix-8 passes and the
    register-resident one-wave passes (radices 16, 16, 4; forced by
    cgm_set_radix_plan), in
    batches of 1, 4 and 16 vectors with their own maps, all three modes,
    against the same product formed by torch.fft in complex128.  Also prints
    the autotuner's own choice."""
    import numpy as np
    import torch
    from cugmres_amd import problems, solver
    N = 512
    ctx = solver.Context(N, 0)
    print("radix plan", ctx.radix_plan(), flush=True)
    ctx.set_radix_plan(wide)
    g = ctx.spectrum_helmholtz(problems.KH)
    blk = ctx.block_helmholtz(problems.KH)
    for B in (1, 4, 16):
        print(f"batched matvec N={N} B={B} wide={wide}", flush=True)
        chis = np.stack([problems.config_plain(N, seed=70 + k)[0].reshape(-1) for k in range(B)])
        c32 = ctx.to_device(chis, "f32").reshape(B, N * N)
        rng = np.random.default_rng(B)
        x32 = torch.as_tensor((rng.standard_normal((B, N * N)) + 1j * rng.standard_normal((B, N * N)))
                              .astype(np.complex64)).to(ctx.device)
        for mode in (0, 1, 2):
            y = ctx.matvec_batched(g, c32, x32, mode).cpu().numpy()
            for k in (0, B - 1):
                ref = _torch_matvec(torch, blk, c32[k].to(torch.float64), x32[k].to(torch.complex128), N, ctx.M,
                                    mode)
                err = rel(y[k], ref.cpu().numpy())
                assert err < 2e-6, (B, mode, k, err)
    ctx.close()

