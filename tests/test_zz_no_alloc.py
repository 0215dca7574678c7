"""cgm_matvec_batched_noalloc: the allocation-free batched matvec against
cgm_matvec_batched (same pass kernels, the FFT intermediate in the library's
fixed device array instead of a context buffer): bit-identical outputs, on
two calls in a row, the second with a smaller batch, and a clear error
beyond the array's size."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _bind():
    from cugmres_amd._lib import lib
    return lib()   # argument and result types from the library's declared signatures


def test_noalloc_matvec_equals_batched_matvec():
    import torch
    from cugmres_amd import problems, solver
    L = _bind()
    N = 256
    ctx = solver.Context(N, 0)
    g = ctx.spectrum_helmholtz(problems.KH)
    rng = np.random.default_rng(7)
    for batch, shared in ((6, False), (3, True)):     # second call smaller
        print(f"noalloc matvec N={N} batch={batch} shared={shared}", flush=True)
        chis = np.stack([problems.config_plain(N, seed=50 + k)[0].reshape(-1) for k in range(batch)])
        dchi = ctx.to_device(chis[0] if shared else chis, "f32")
        if not shared:
            dchi = dchi.reshape(batch, N * N)
        x = torch.as_tensor((rng.standard_normal((batch, N * N)) + 1j * rng.standard_normal((batch, N * N)))
                            .astype(np.complex64)).to(ctx.device)
        for mode in (0, 1, 2):
            ref = ctx.matvec_batched(g, dchi, x, mode)
            y = torch.full_like(x, float("nan"))
            st = L.cgm_matvec_batched_noalloc(ctx.handle, ctypes.c_void_p(g.data_ptr()),
                                              ctypes.c_void_p(dchi.data_ptr()), 0 if shared else N * N,
                                              ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(y.data_ptr()),
                                              batch, mode)
            assert st == 0
            torch.cuda.synchronize()
            assert torch.equal(y, ref), (batch, mode)
    # beyond the fixed array: a clear error, nothing launched
    cap = L.cgm_noalloc_scratch_bytes()
    too_many = cap // (8 * N * ctx.M) + 1
    st = L.cgm_matvec_batched_noalloc(ctx.handle, ctypes.c_void_p(g.data_ptr()), None, 0,
                                      ctypes.c_void_p(x.data_ptr()), ctypes.c_void_p(x.data_ptr()), too_many, 2)
    assert st == 1
    assert b"FFT scratch" in L.cgm_last_error()
    ctx.close()
