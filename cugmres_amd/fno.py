"""Learned right preconditioner: a Fourier neural operator (FNO) evaluated and
trained by the library's BF16 MFMA kernels (cgm_fno_* in include/cugmres.h).

The reference has no preconditioner; this is the learned extension that
cgm_fgmres applies to each Krylov vector.  For a unit-norm vector v on the
N x N grid and contrast map chi:

    P(v) = s_id v + FNO(chi * chi_scale, N Re v, N Im v) / N

with a 3 -> 32 lift, four spectral layers (32 channels, retained modes
kx in {0..15, N-16..N-1}, ky in {0..15}, GELU between layers) and a 32 -> 2
projection; s_id is a learned weight of the identity path.  Parameters are
one flat float32 array; `layout()` gives the
offset and shape of every block in the order the C library expects.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, lib

C = 32
KX = 32
KY = 16
LAYERS = 4
LAYER_PARAMS = C * C + C + KX * KY * C * C * 2
NUM_PARAMS = 4 * C + LAYERS * LAYER_PARAMS + 2 * C + 3
CHI_SCALE = 1.0 / 11.1   # maps the binary maps' chi in {0, 11.1} to {0, 1}


def layout():
    """name -> (offset, shape) of every parameter block (float32 units)."""
    out = {"W_in": (0, (3, C)), "b_in": (3 * C, (C,))}
    for l in range(LAYERS):
        base = 4 * C + l * LAYER_PARAMS
        out[f"W{l}"] = (base, (C, C))
        out[f"b{l}"] = (base + C * C, (C,))
        out[f"S{l}"] = (base + C * C + C, (KX, KY, C, C, 2))
    base = 4 * C + LAYERS * LAYER_PARAMS
    out["W_out"] = (base, (C, 2))
    out["b_out"] = (base + 2 * C, (2,))
    out["s_id"] = (base + 2 * C + 2, (1,))
    return out


def unpack(params):
    """Views of a flat parameter array (numpy or torch) by block name."""
    res = {}
    for name, (off, shape) in layout().items():
        size = int(np.prod(shape))
        res[name] = params[off:off + size].reshape(shape)
    return res


def init_params(seed=0, zero_out=True):
    """Seeded initial parameters (float32).  Dense weights are N(0, 1/fan_in),
    spectral weights (re, im) uniform in [0, 1/(C*C)) as in Li et al.'s FNO,
    biases zero, s_id = 1.  zero_out=True zeroes W_out so P starts as the
    identity."""
    rng = np.random.default_rng(seed)
    p = np.zeros(NUM_PARAMS, np.float32)
    v = unpack(p)
    v["W_in"][...] = rng.normal(0.0, 1.0 / np.sqrt(3.0), (3, C))
    for l in range(LAYERS):
        v[f"W{l}"][...] = rng.normal(0.0, 1.0 / np.sqrt(C), (C, C))
        v[f"S{l}"][...] = rng.uniform(0.0, 1.0 / (C * C), (KX, KY, C, C, 2))
    if not zero_out:
        v["W_out"][...] = rng.normal(0.0, 1.0 / np.sqrt(C), (C, 2))
    v["s_id"][...] = 1.0
    return p


def save(path, params, **meta):
    np.savez(path, params=np.asarray(params, np.float32),
             **{k: np.asarray(val) for k, val in meta.items()})


def load(path):
    with np.load(path, allow_pickle=False) as f:
        p = np.asarray(f["params"], np.float32)
    if p.size != NUM_PARAMS:
        raise ValueError(f"{path}: {p.size} parameters, expected {NUM_PARAMS}")
    return p


class Preconditioner:
    """The network on one solver context (cugmres_amd.solver.Context).

    `batch` is the largest number of vectors one forward call takes.  The
    parameters live in a float32 tensor on the context's device; the C
    library keeps a reference to it, so this object owns it."""

    def __init__(self, ctx, batch=1, chi_scale=CHI_SCALE, params=None):
        self.ctx = ctx
        self.torch = ctx.torch
        self.batch = int(batch)
        n = ctypes.c_size_t()
        check(lib().cgm_fno_num_params(ctypes.byref(n)), "cgm_fno_num_params")
        if n.value != NUM_PARAMS:
            raise RuntimeError(f"library has {n.value} FNO parameters, module expects {NUM_PARAMS}")
        h = ctypes.c_void_p()
        check(lib().cgm_fno_create(ctx.handle, self.batch, float(chi_scale), ctypes.byref(h)), "cgm_fno_create")
        self._h = h
        self.params = None
        self.set_params(init_params() if params is None else params)

    @property
    def handle(self):
        return self._h

    def close(self):
        if getattr(self, "_h", None) is not None:
            lib().cgm_fno_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_params(self, params):
        t = self.torch
        if isinstance(params, np.ndarray):
            params = t.from_numpy(np.ascontiguousarray(params, np.float32))
        params = params.to(device=self.ctx.device, dtype=t.float32).contiguous()
        if params.numel() != NUM_PARAMS:
            raise ValueError(f"{params.numel()} parameters, expected {NUM_PARAMS}")
        self.params = params
        check(lib().cgm_fno_set_params(self._h, ctypes.c_void_p(params.data_ptr())), "cgm_fno_set_params")

    def forward(self, chi, v, train=False):
        """z = P(v) for a batch: chi (B, N*N) float32 and v (B, N*N) complex64
        (unit norm per row) on the device."""
        t = self.torch
        B = v.shape[0] if v.dim() == 2 else 1
        z = t.empty_like(v)
        check(lib().cgm_fno_forward(self._h, ctypes.c_void_p(chi.data_ptr()), ctypes.c_void_p(v.data_ptr()),
                                    ctypes.c_void_p(z.data_ptr()), int(B), int(bool(train))), "cgm_fno_forward")
        return z

    def backward(self, dz, grad=None):
        """Parameter gradient of a real loss given dz = dL/dRe z + i dL/dIm z
        of the last forward(train=True)."""
        t = self.torch
        if grad is None:
            grad = t.empty(NUM_PARAMS, dtype=t.float32, device=self.ctx.device)
        check(lib().cgm_fno_backward(self._h, ctypes.c_void_p(dz.data_ptr()), ctypes.c_void_p(grad.data_ptr())),
              "cgm_fno_backward")
        return grad
