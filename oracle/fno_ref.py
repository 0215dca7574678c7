"""CPU float64 restatement of the FNO preconditioner (cugmres_amd/csrc/cgm_fno.hip).

TEST INFRASTRUCTURE ONLY: imported by tests/ as the checker of the library's
BF16 MFMA forward and backward kernels.  Nothing in ``cugmres_amd`` imports it.

The reference has no preconditioner, so there is no reference file to follow:
the definition is the one in include/cugmres.h and cugmres_amd/fno.py, written
here with torch's CPU FFTs (rfft2 / irfft2 of the truncated spectrum) and
torch autograd in float64 for the gradients.  The tests pin this restatement
against explicit DFT sums, a closed-form single-mode response, the identity
at W_out = 0 and finite differences.
"""
import math

import numpy as np
import torch

C = 32
KX = 32
KY = 16
LAYERS = 4
LAYER_PARAMS = C * C + C + KX * KY * C * C * 2
NUM_PARAMS = 4 * C + LAYERS * LAYER_PARAMS + 2 * C + 3


def blocks(p):
    """Views of a flat float64 torch parameter vector by block name."""
    out = {}
    out["W_in"] = p[0:3 * C].reshape(3, C)
    out["b_in"] = p[3 * C:4 * C]
    for l in range(LAYERS):
        base = 4 * C + l * LAYER_PARAMS
        out[f"W{l}"] = p[base:base + C * C].reshape(C, C)
        out[f"b{l}"] = p[base + C * C:base + C * C + C]
        s = p[base + C * C + C:base + LAYER_PARAMS].reshape(KX, KY, C, C, 2)
        out[f"S{l}"] = torch.complex(s[..., 0], s[..., 1])
    base = 4 * C + LAYERS * LAYER_PARAMS
    out["W_out"] = p[base:base + 2 * C].reshape(C, 2)
    out["b_out"] = p[base + 2 * C:base + 2 * C + 2]
    out["s_id"] = p[base + 2 * C + 2]
    return out


def kx_index(N):
    return list(range(16)) + list(range(N - 16, N))


def spectral_conv(h, S):
    """K(h) for h (B, N, N, C) real: rfft2 over (x, y), mixing of the retained
    modes by S (KX, KY, C_in, C_out), irfft2 of the truncated spectrum."""
    B, N = h.shape[0], h.shape[1]
    hf = torch.fft.rfft2(h.permute(0, 3, 1, 2))               # (B, C, N, N/2+1)
    idx = kx_index(N)
    sel = hf[:, :, idx, :KY]                                  # (B, C, KX, KY)
    mixed = torch.einsum("bckq,kqco->bokq", sel, S)
    out = torch.zeros(B, S.shape[3], N, N // 2 + 1, dtype=hf.dtype)
    out[:, :, idx, :KY] = mixed
    return torch.fft.irfft2(out, s=(N, N)).permute(0, 2, 3, 1)


def gelu(x):
    return torch.nn.functional.gelu(x, approximate="tanh")


def forward(p, chi, v, chi_scale):
    """z = s_id v + FNO(chi * chi_scale, N Re v, N Im v) / N.
    p: float64 (NUM_PARAMS,), chi: (B, N, N) float64, v: (B, N, N) complex128."""
    w = blocks(p)
    N = chi.shape[-1]
    a = torch.stack([chi * chi_scale, N * v.real, N * v.imag], dim=-1)
    h = a @ w["W_in"] + w["b_in"]
    for l in range(LAYERS):
        pre = h @ w[f"W{l}"] + w[f"b{l}"] + spectral_conv(h, w[f"S{l}"])
        h = gelu(pre) if l < LAYERS - 1 else pre
    out = h @ w["W_out"] + w["b_out"]
    return w["s_id"] * v + torch.complex(out[..., 0], out[..., 1]) / N


def forward_np(params, chi, v, chi_scale):
    p = torch.from_numpy(np.asarray(params, np.float64))
    z = forward(p, torch.from_numpy(np.asarray(chi, np.float64)),
                torch.from_numpy(np.asarray(v, np.complex128)), chi_scale)
    return z.numpy()


def grad_np(params, chi, v, chi_scale, dz):
    """Gradient of L = sum(Re dz * Re z + Im dz * Im z) with respect to the parameters."""
    p = torch.tensor(np.asarray(params, np.float64), requires_grad=True)
    z = forward(p, torch.from_numpy(np.asarray(chi, np.float64)),
                torch.from_numpy(np.asarray(v, np.complex128)), chi_scale)
    dzt = torch.from_numpy(np.asarray(dz, np.complex128))
    L = (z.real * dzt.real + z.imag * dzt.imag).sum()
    L.backward()
    return p.grad.numpy()


def spectral_conv_dense(h, S):
    """The same K(h) by explicit DFT sums (O(N^4)); checks spectral_conv."""
    B, N = h.shape[0], h.shape[1]
    x = torch.arange(N, dtype=torch.float64)
    kxs = torch.tensor(kx_index(N), dtype=torch.float64)
    kys = torch.arange(KY, dtype=torch.float64)
    ex = torch.exp(-2j * math.pi * torch.outer(kxs, x) / N)          # (KX, N)
    ey = torch.exp(-2j * math.pi * torch.outer(kys, x) / N)          # (KY, N)
    hc = h.to(torch.complex128)
    U = torch.einsum("kx,qy,bxyc->bckq", ex, ey, hc)
    V = torch.einsum("bckq,kqco->bokq", U, S)
    w = torch.tensor([1.0] + [2.0] * (KY - 1), dtype=torch.float64)
    val = torch.einsum("kx,qy,bokq,q->bxyo", ex.conj(), ey.conj(), V, w.to(torch.complex128))
    return val.real / (N * N)
