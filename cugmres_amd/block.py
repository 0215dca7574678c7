"""Block GMRES over the right-hand sides of one contrast map.

Reference counterpart: the one-system restarted GMRES of
B/CUDA_GMRES/GMRES.cuh:22-446, whose Arnoldi step projects one vector at a
time (the MGS loop of dots at GMRES.cuh:289-303).  Here the s right-hand sides
of one map (the batched workload's plane waves at s angles) share one block
Krylov space K_k(A P, R0), right-preconditioned by the map's tile inverses P:

  * every step applies the operator and P to s vectors at once
    (cgm_matvec_batched, cgm_tiles_apply on the device);
  * the block classical Gram-Schmidt projection H = V^H W and the update
    W -= V H (twice: CGS2) are GEMMs of (k s) x n by n x s, which rocBLAS runs
    on the matrix cores (FP32 MFMA), instead of one GEMV per system: the
    projection as a batched complex GEMM over chunks of n, the update as one
    real GEMM over the basis seen as real rows;
  * W = V' S is orthonormalised by CholQR2 (Gram matrix and Cholesky factor in
    FP64);
  * the block Hessenberg matrix is reduced by one 2s x s Householder QR per
    step in FP64 on the host, which gives every column's residual norm after
    each step; a cycle ends when every column is below the tolerance or the
    block breaks down (a rank-deficient W), and the next cycle restarts from
    the true residual of the columns still above it.

Vectors are rows: B and X are (s, n), the basis is ((m + 1) s, n).  The core
`block_gmres` takes the operator and the preconditioner as callables on such
row blocks, so it runs on CPU tensors as well as on the device;
`solve_map` wires it to a Context and a Tiles object."""
import numpy as np


class BlockBreakdown(Exception):
    """W has (numerically) dependent columns: the block Krylov space stopped growing."""


def inner(torch, A, B, chunk=4096):
    """conj(A) B^T for row blocks A (a, n), B (b, n): the (a, b) matrix of
    inner products <A_i, B_j>.  n is split into chunks of `chunk` so that the
    product is a batched GEMM with n / chunk members (summed afterwards)
    rather than one GEMM with an n-long inner dimension and an a x b output,
    which the BLAS spreads over a handful of work-groups."""
    a, n = A.shape
    b = B.shape[0]
    c = n // chunk if (chunk and n % chunk == 0 and n > chunk) else 1
    Ac = A.conj().resolve_conj().reshape(a, c, n // c).transpose(0, 1)    # (c, a, n/c)
    Bc = B.reshape(b, c, n // c).transpose(0, 1)                           # (c, b, n/c)
    return torch.bmm(Ac, Bc.transpose(1, 2)).sum(0)


def project(torch, Vk, W, chunk=4096):
    """V^H W, (k, s), for the basis rows Vk (k, n) and the block W (s, n): per
    chunk of n the product V_chunk conj(W_chunk)^T, one batched GEMM whose
    large side is the basis (measured faster than conj(W) V^T, whose k-wide
    output sits on the BLAS tile's other side), summed over the chunks and
    conjugated.  Only the s rows of W are conjugated, never the basis."""
    k, n = Vk.shape
    s = W.shape[0]
    c = n // chunk if (chunk and n % chunk == 0 and n > chunk) else 1
    Wc = W.conj().resolve_conj().reshape(s, c, n // c).transpose(0, 1)    # (c, s, n/c)
    Vc = Vk.reshape(k, c, n // c).transpose(0, 1)                          # (c, k, n/c)
    return torch.bmm(Vc, Wc.transpose(1, 2)).sum(0).conj()


def update(torch, W, Vk, H):
    """W - (V H)^T for H (k, s): one real GEMM of [Re H^T; Im H^T] (2s x k) by
    the basis seen as real rows (k x 2n, real and imaginary parts
    interleaved), so the basis is read once; the four real products combine
    into the complex one.  (A complex GEMM with s = 16 output rows runs the
    BLAS's 128-row matrix-core tiles mostly empty.)"""
    s, n = W.shape
    k = Vk.shape[0]
    Vr = torch.view_as_real(Vk).reshape(k, 2 * n)
    HH = torch.cat([H.real.T, H.imag.T]).contiguous()
    C = (HH @ Vr).reshape(2, s, n, 2)
    return W - torch.complex(C[0, ..., 0] - C[1, ..., 1], C[0, ..., 1] + C[1, ..., 0])


def _cholqr(torch, W, tol=1e-12):
    """Rows W (s, n) = conj(L) V with V's rows orthonormal: returns V (s, n)
    complex128 and S = L^H (s, s), so that the columns satisfy
    W^T = V^T S.  Raises BlockBreakdown when the Gram matrix is not safely
    positive definite."""
    Wd = W.to(torch.complex128)
    G = inner(torch, Wd, Wd, chunk=1024)      # G[a, b] = <W_a, W_b>, Hermitian
    G = 0.5 * (G + G.mH)
    d = torch.real(torch.diagonal(G)).clamp_min(0.0)
    L, info = torch.linalg.cholesky_ex(G)
    if int(info) != 0:
        raise BlockBreakdown("Gram matrix not positive definite")
    ld = torch.abs(torch.diagonal(L)) ** 2
    if bool((ld <= tol * d.max()).any()):
        raise BlockBreakdown("dependent columns")
    V = torch.linalg.inv(L.conj()) @ Wd      # s x s inverse: a GEMM over n, not a triangular solve
    return V, L.conj().T


def cholqr2(torch, W):
    """CholQR twice: W^T = V^T S with V orthonormal to working precision."""
    V1, S1 = _cholqr(torch, W)
    V, S2 = _cholqr(torch, V1)
    return V, S2 @ S1


class _HessQR:
    """Householder QR of a block upper-Hessenberg matrix built one block column
    (height (j + 2) s, width s) at a time, and the rotated right-hand side
    G = Q^H [S0; 0] (FP64, numpy)."""

    def __init__(self, S0, m):
        s = S0.shape[0]
        self.s = s
        self.R = np.zeros(((m + 1) * s, m * s), dtype=np.complex128)
        self.G = np.zeros(((m + 1) * s, s), dtype=np.complex128)
        self.G[:s] = S0
        self.Qh = []   # per step: the 2s x 2s unitary Q_j^H acting on rows j s .. (j + 2) s
        self.k = 0

    def add(self, Hcol):
        """Hcol: ((k + 2) s, s) block column k of the Hessenberg matrix.
        Returns the residual norm of every right-hand side after k + 1 steps."""
        s, k = self.s, self.k
        h = np.array(Hcol, dtype=np.complex128, copy=True)
        for j, Qh in enumerate(self.Qh):
            h[j * s:(j + 2) * s] = Qh @ h[j * s:(j + 2) * s]
        Q, Rb = np.linalg.qr(h[k * s:(k + 2) * s], mode="complete")
        Qh = Q.conj().T
        self.Qh.append(Qh)
        h[k * s:(k + 2) * s] = 0.0
        h[k * s:(k + 1) * s] = Rb[:s]
        self.R[:(k + 2) * s, k * s:(k + 1) * s] = h
        self.G[k * s:(k + 2) * s] = Qh @ self.G[k * s:(k + 2) * s]
        self.k = k + 1
        return np.linalg.norm(self.G[(k + 1) * s:(k + 2) * s], axis=0)

    def solve(self):
        """Y ((k s) x s) minimising || [S0; 0] - Hbar Y || column by column."""
        n = self.k * self.s
        if n == 0:
            return np.zeros((0, self.s), dtype=np.complex128)
        import scipy.linalg
        return scipy.linalg.solve_triangular(self.R[:n, :n], self.G[:n], lower=False)


def block_gmres(torch, apply_A, apply_P, B, X=None, restart=64, tol=1e-6, max_cycles=20, residual=None,
                on_step=None, stall_window=10, stall_ratio=0.9):
    """Right-preconditioned block GMRES(restart) for A X_c = B_c, every row c
    of B (s, n) a right-hand side; restart counts block steps.

    apply_A(Y) and apply_P(Y) map a row block (r, n) to (r, n) in B's dtype;
    X (optional) may be of a wider dtype (complex128), and the corrections are
    added to it in that precision; residual(X) (optional) returns B - A X in
    any higher precision for the convergence test and the restarts (default:
    B - apply_A(X)).  Each cycle restarts from that residual, so with a
    complex128 X and an FP64 residual the cycles are iterative refinement
    steps and reach tolerances below the working precision's floor.  Inside a
    cycle the estimates are driven to tol / 2.  A cycle also ends at the
    working precision's floor: when, with every estimate within 20 x tol and
    1e-4 times the cycle's starting residual, the largest one has not fallen
    below stall_ratio times its value stall_window steps before.  Returns (X, info)
    with info["rel_residual"] (per row, from residual()), info["converged"]
    (per row), info["block_steps"], info["cycles"] and info["breakdowns"]."""
    s, n = B.shape
    dev, dt = B.device, B.dtype
    if X is None:
        X = torch.zeros_like(B)
    resid = residual if residual is not None else (lambda Xc: B - apply_A(Xc))
    bnorm = torch.linalg.vector_norm(B.to(torch.complex128), dim=1).cpu().numpy()
    bnorm = np.where(bnorm > 0, bnorm, 1.0)
    Vb = torch.empty(((restart + 1) * s, n), dtype=dt, device=dev)
    on_device = dev.type == "cuda"
    steps = cycles = breaks = 0
    R = resid(X)
    rel = torch.linalg.vector_norm(R.to(torch.complex128), dim=1).cpu().numpy() / bnorm
    for _ in range(max_cycles):
        active = np.nonzero(rel > tol)[0]
        if active.size == 0:
            break
        cycles += 1
        idx = torch.as_tensor(active, device=dev)
        Ra = R.index_select(0, idx).to(dt)
        sa = int(active.size)
        try:
            V0, S0 = cholqr2(torch, Ra)
        except BlockBreakdown:
            # dependent residuals: this cycle takes the first column alone, the
            # others follow in later cycles from their updated residuals
            sa = 1
            idx = idx[:1]
            Ra = Ra[:1]
            V0, S0 = cholqr2(torch, Ra)
            active = active[:1]
        Vb[:sa] = V0.to(dt)
        qr = _HessQR(S0.cpu().numpy(), restart)
        bn = bnorm[active]
        start = float(rel[active].max())
        hist = []
        W_next = None
        for j in range(restart):
            k = (j + 1) * sa
            Vk = Vb[:k]
            W = W_next if W_next is not None else apply_A(apply_P(Vb[j * sa:k]))
            W_next = None
            H1 = project(torch, Vk, W)                              # (k, sa) = V^H W
            W = update(torch, W, Vk, H1)
            H2 = project(torch, Vk, W)
            W = update(torch, W, Vk, H2)
            Hc = (H1 + H2).to(torch.complex128)
            steps += 1
            try:
                Vn, Sj = cholqr2(torch, W)
            except BlockBreakdown:
                breaks += 1
                Hcol = torch.cat([Hc, torch.zeros((sa, sa), dtype=torch.complex128, device=dev)]).cpu().numpy()
                qr.add(Hcol)
                break
            Vb[k:k + sa] = Vn.to(dt)
            Hd = torch.cat([Hc, Sj])
            if on_device:
                # the Hessenberg column leaves through a pinned copy; the next
                # step's operator and tile inverses are queued behind it, so
                # the device computes them while the host updates the QR
                Hh = torch.empty(Hd.shape, dtype=Hd.dtype, pin_memory=True)
                Hh.copy_(Hd, non_blocking=True)
                ev = torch.cuda.Event()
                ev.record()
                if j + 1 < restart:
                    W_next = apply_A(apply_P(Vb[k:k + sa]))
                ev.synchronize()
                Hcol = Hh.numpy()
            else:
                Hcol = Hd.cpu().numpy()
            est = qr.add(Hcol) / bn
            if on_step is not None:
                on_step(steps, est)
            if bool((est <= 0.5 * tol).all()):
                break
            hist.append(float(est.max()))
            if (len(hist) > stall_window and hist[-1] <= 20 * tol and hist[-1] <= 1e-4 * start
                    and hist[-1] > stall_ratio * hist[-1 - stall_window]):
                break
        Y = torch.as_tensor(qr.solve(), device=dev)                  # (qr.k sa, sa)
        U = (Y.T.to(dt) @ Vb[:qr.k * sa])                            # (sa, n): (V Y)^T
        Xa = X.index_select(0, idx) + apply_P(U).to(X.dtype)
        X = X.index_copy(0, idx, Xa.to(X.dtype))
        R = resid(X)
        rel = torch.linalg.vector_norm(R.to(torch.complex128), dim=1).cpu().numpy() / bnorm
    info = {"rel_residual": rel.tolist(), "converged": (rel <= tol).tolist(), "block_steps": steps,
            "cycles": cycles, "breakdowns": breaks}
    return X, info


def solve_map(ctx, gamma, chi, B, tiles=None, X=None, restart=64, tol=1e-6, max_cycles=20, gamma64=None,
              chi64=None, on_step=None):
    """Block GMRES for the right-hand sides B (s, N*N) complex64 of one map chi
    (N*N,) float32 on the device, preconditioned by `tiles` (a factored Tiles
    object of that map; None = no preconditioner).  With gamma64 and chi64
    (the FP64 spectrum and map) the solution is kept in complex128 and the
    restarts and the convergence test use the FP64 true residual
    (cgm_matvec_c128): the cycles then refine in FP64 past the complex64 floor.
    Else everything is complex64.  Returns (X, info) as block_gmres."""
    torch = ctx.torch
    from . import _lib
    n = ctx.N * ctx.N
    if B.dim() != 2 or B.shape[1] != n or B.dtype != torch.complex64:
        raise ValueError(f"solve_map: B must be (s, {n}) complex64, got {tuple(B.shape)} {B.dtype}")
    if chi.dim() != 1 or chi.shape[0] != n or chi.dtype != torch.float32:
        raise ValueError(f"solve_map: chi must be ({n},) float32 (one map), got {tuple(chi.shape)} {chi.dtype}")
    if (gamma64 is None) != (chi64 is None):
        raise ValueError("solve_map: give both gamma64 and chi64 (the FP64 path) or neither")
    s = int(B.shape[0])
    if tiles is not None and tiles.maps != 1:
        raise ValueError("solve_map: the tiles must be factored for one map")

    def apply_A(Y):
        return ctx.matvec_batched(gamma, chi, Y.contiguous(), _lib.OP_FORWARD)

    def apply_P(Y):
        return tiles.apply(Y.contiguous()) if tiles is not None else Y

    residual = None
    if gamma64 is not None and chi64 is not None:
        B64 = B.to(torch.complex128)
        X = torch.zeros_like(B64) if X is None else X.to(torch.complex128)

        def residual(Xc):
            X64 = Xc.to(torch.complex128)
            return torch.stack([B64[c] - ctx.matvec_c128(gamma64, chi64, X64[c].contiguous(), _lib.OP_FORWARD)
                                for c in range(s)])
    return block_gmres(torch, apply_A, apply_P, B, X=X, restart=restart, tol=tol, max_cycles=max_cycles,
                       residual=residual, on_step=on_step)
