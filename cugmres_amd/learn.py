"""Training of the learned right preconditioner on the GPU.

The network (cugmres_amd.fno) is trained by the library's own BF16 MFMA
forward and backward kernels; torch provides the device buffers, the sample
generators and the Adam update.  The loss is the preconditioned residual of
unit-norm vectors v,

    L = mean_b || A P(v_b) - v_b ||^2,      dL/dz_b = 2 A^H (A z_b - v_b) / B,

formed with the plain solver's own operator (cgm_matvec; A^H y =
conj(A^T conj(y)) from the adjoint mode), so no converged solutions are
needed as targets: the operator supervises every sample exactly.  Samples are
seeded (sample k of a run uses seed base + k): permittivity maps of the plain
workload's kind and vectors drawn from what flexible GMRES hands the
preconditioner - incident plane waves (the first Krylov vector), their images
under A, and band-limited random fields (later residual directions).
"""
import argparse
import math
import time

import numpy as np

from . import fno, problems
from ._lib import OP_ADJOINT, OP_FORWARD


def sample_maps(N, seeds, block=8, fill=0.75):
    """Binary {1, 12.1} maps of block x block cells over the central `fill` of
    the area (the plain workload's kind), chi = eps - 1."""
    out = np.empty((len(seeds), N, N), np.float32)
    for k, s in enumerate(seeds):
        out[k] = problems.block_eps_map(N, block, 0.5, fill, seed=int(s)) - 1.0
    return out


class Sampler:
    """Seeded training vectors on the device (complex64, unit norm)."""

    def __init__(self, ctx, gamma, kh=problems.KH):
        self.ctx = ctx
        self.t = ctx.torch
        self.g = gamma
        self.kh = kh
        N = ctx.N
        t = self.t
        self.i = t.arange(N, dtype=t.float64, device=ctx.device)[:, None]
        self.j = t.arange(N, dtype=t.float64, device=ctx.device)[None, :]
        f = t.fft.fftfreq(N, device=ctx.device, dtype=t.float64)
        self.f2 = (f[:, None] ** 2 + f[None, :] ** 2)

    def plane(self, theta):
        t = self.t
        ph = -self.kh * (self.i * math.cos(theta) + self.j * math.sin(theta))
        return t.polar(t.ones_like(ph), ph).to(t.complex64).reshape(-1)

    def smooth(self, gen, ell):
        """Gaussian-filtered complex noise with correlation length ell cells."""
        t = self.t
        N = self.ctx.N
        w = t.randn(N, N, dtype=t.float64, device=self.ctx.device, generator=gen) + \
            1j * t.randn(N, N, dtype=t.float64, device=self.ctx.device, generator=gen)
        filt = t.exp(-2.0 * (math.pi * ell) ** 2 * self.f2)
        return t.fft.ifft2(t.fft.fft2(w) * filt).to(t.complex64).reshape(-1)

    def batch(self, chi_d, seed):
        """One vector per map in chi_d (B, N*N): kind k = seed % 4 picks a plane
        wave (k = 0, 1), A times a plane wave (k = 2) or a smooth field (k = 3)."""
        t = self.t
        B = chi_d.shape[0]
        out = t.empty(B, self.ctx.N * self.ctx.N, dtype=t.complex64, device=self.ctx.device)
        for b in range(B):
            s = int(seed) + b
            rng = np.random.default_rng(s)
            theta = 2 * math.pi * rng.uniform()
            kind = s % 4
            if kind <= 1:
                v = self.plane(theta)
            elif kind == 2:
                v = self.ctx.matvec(self.g, chi_d[b], self.plane(theta), OP_FORWARD)
            else:
                gen = t.Generator(device=self.ctx.device)
                gen.manual_seed(s)
                v = self.smooth(gen, float(rng.uniform(0.5, 8.0)))
            out[b] = v / t.linalg.vector_norm(v)
        return out


def residual_chain(ctx, gamma, net, chi_d, v, depth, tiles=None):
    """depth steps of v <- (v - A P(v)) / ||v - A P(v)|| with the current
    preconditioner (the network, or the tile inverses followed by the
    network): the residuals of preconditioned Richardson iteration, which span
    the Krylov space flexible GMRES builds with P, so training on them aims
    the loss at the directions P still fails on."""
    t = ctx.torch
    from .tiles import TwoLevel
    P = TwoLevel(tiles, net) if tiles is not None else None
    for _ in range(depth):
        z = P.apply(ctx, gamma, chi_d, v) if P else net.forward(chi_d, v)
        for b in range(v.shape[0]):
            r = v[b] - ctx.matvec(gamma, chi_d[b], z[b], OP_FORWARD)
            v[b] = r / t.linalg.vector_norm(r)
    return v


def residual_loss_grad(ctx, gamma, chi_d, v, z):
    """(L, dL/dz) for L = mean_b ||A z_b - v_b||^2."""
    t = ctx.torch
    B = v.shape[0]
    dz = t.empty_like(z)
    L = 0.0
    for b in range(B):
        r = ctx.matvec(gamma, chi_d[b], z[b], OP_FORWARD) - v[b]
        L += float(t.sum(r.real ** 2 + r.imag ** 2))
        ah = t.conj_physical(ctx.matvec(gamma, chi_d[b], t.conj_physical(r), OP_ADJOINT))
        dz[b] = ah * (2.0 / B)
    return L / B, dz


class Adam:
    """Adam (Kingma & Ba, 2015) on a flat float32 device tensor."""

    def __init__(self, p, lr=1e-3, b1=0.9, b2=0.999, eps=1e-8):
        t = p
        self.m = t.new_zeros(t.shape)
        self.v = t.new_zeros(t.shape)
        self.lr, self.b1, self.b2, self.eps = lr, b1, b2, eps
        self.k = 0

    def step(self, p, g, lr=None):
        self.k += 1
        lr = self.lr if lr is None else lr
        self.m.mul_(self.b1).add_(g, alpha=1 - self.b1)
        self.v.mul_(self.b2).addcmul_(g, g, value=1 - self.b2)
        mh = self.m / (1 - self.b1 ** self.k)
        vh = self.v / (1 - self.b2 ** self.k)
        p.addcdiv_(mh, vh.sqrt_().add_(self.eps), value=-lr)


def validation_set(ctx, gamma, batch, tiles_T=0, seed=0, chain=3):
    """Held-out (maps, vectors) for validation: maps and vectors from seeds
    that training never draws (2e9 + ...).  With tiles_T the vectors are the
    normalised residuals the tile inverses leave, sample b taken b % (chain + 1)
    steps down the Richardson chain of the tile inverses alone (no network),
    so the set stays fixed while the network trains.  Returns (chi, v, tiles)."""
    t = ctx.torch
    seeds = [2_000_000_000 + seed * 1_000_003 + b for b in range(batch)]
    chi_d = t.from_numpy(sample_maps(ctx.N, seeds).reshape(batch, -1)).to(ctx.device)
    v = Sampler(ctx, gamma).batch(chi_d, seeds[0])
    tl = None
    if tiles_T:
        from .tiles import Tiles, tile_residual
        tl = Tiles(ctx, tiles_T, batch).factor(ctx.block_helmholtz(problems.KH), chi_d)
        for d in range(chain):
            r = tile_residual(ctx, gamma, tl, chi_d, v.contiguous())[1]
            for b in range(batch):
                if b % (chain + 1) > d:
                    v[b] = r[b]
        v = tile_residual(ctx, gamma, tl, chi_d, v.contiguous())[1].contiguous()
    return chi_d, v, tl


def validation_loss(ctx, gamma, net, chi_d, v):
    z = net.forward(chi_d, v)
    return residual_loss_grad(ctx, gamma, chi_d, v, z)[0]


def train(ctx, samples=20000, batch=8, lr=1e-3, seed=0, params=None, log=None, log_every=100,
          clip=1.0, identity_weight=0.0, chain=3, tiles_T=0, map_pool=8, val_every=0):
    """Trains a Preconditioner on `samples` seeded (map, vector) pairs at the
    context's N, starting from `params` or from init_params(seed) with the
    identity path weighted by identity_weight.  Step k trains on vectors
    taken k % (chain + 1) steps down the residual chain.  With tiles_T > 0
    the network is the second level of tiles.TwoLevel and trains on the
    normalised residuals r = v - A T v that the T x T tile inverses of its
    maps leave; the maps then come from a pool of map_pool batches whose
    inverses are factored once (step k uses batch k % map_pool; 0 = new maps
    factored at every step), while the vectors stay new at every step.
    With val_every > 0 the loss on a fixed held-out set (validation_set) is
    evaluated every val_every steps and the parameters of its lowest value are
    the ones returned (net.validation holds the curve, net.best_step the step).
    Returns (Preconditioner, history of (step, loss))."""
    t = ctx.torch
    gamma = ctx.spectrum_helmholtz(problems.KH)
    if params is None:
        p0 = fno.init_params(seed, zero_out=True)
        fno.unpack(p0)["s_id"][...] = identity_weight
    else:
        p0 = np.asarray(params, np.float32)
    net = fno.Preconditioner(ctx, batch=batch, params=p0)
    p = net.params.clone()
    opt = Adam(p, lr)
    sampler = Sampler(ctx, gamma)
    tl = None
    pool = []
    if tiles_T:
        from .tiles import Tiles, tile_residual
        blk = ctx.block_helmholtz(problems.KH)
        for q in range(map_pool):
            seeds = [seed * 1_000_003 + q * batch + b for b in range(batch)]
            chi_q = t.from_numpy(sample_maps(ctx.N, seeds).reshape(batch, -1)).to(ctx.device)
            pool.append((chi_q, Tiles(ctx, tiles_T, batch).factor(blk, chi_q)))
        if not pool:
            tl = Tiles(ctx, tiles_T, batch)
    steps = max(1, samples // batch)
    hist = []
    val = None
    best = (float("inf"), None, -1)
    if val_every:
        val = validation_set(ctx, gamma, batch, tiles_T, seed, chain)
        net.validation = []
    t0 = time.time()
    for k in range(steps):
        seeds = [seed * 1_000_003 + k * batch + b for b in range(batch)]
        if pool:
            chi_d, tl = pool[k % len(pool)]
        else:
            chi_d = t.from_numpy(sample_maps(ctx.N, seeds).reshape(batch, -1)).to(ctx.device)
        v = sampler.batch(chi_d, seeds[0])
        if tiles_T and not pool:
            tl.factor(blk, chi_d)
        if chain:
            v = residual_chain(ctx, gamma, net, chi_d, v, k % (chain + 1), tl)
        if tl is not None:
            v = tile_residual(ctx, gamma, tl, chi_d, v)[1].contiguous()
        z = net.forward(chi_d, v, train=True)
        L, dz = residual_loss_grad(ctx, gamma, chi_d, v, z)
        g = net.backward(dz)
        gn = float(t.linalg.vector_norm(g))
        if not math.isfinite(gn):
            raise FloatingPointError(f"non-finite gradient at step {k}")
        if clip and gn > clip:
            g.mul_(clip / gn)
        cur = lr * 0.5 * (1 + math.cos(math.pi * k / steps))   # cosine decay
        opt.step(p, g, cur)
        net.set_params(p)
        hist.append((k, L))
        if val is not None and ((k + 1) % val_every == 0 or k == steps - 1):
            Lv = validation_loss(ctx, gamma, net, val[0], val[1])
            net.validation.append((k + 1, Lv))
            if Lv < best[0]:
                best = (Lv, p.clone(), k + 1)
        if log and (k % log_every == 0 or k == steps - 1):
            vtxt = f" validation {net.validation[-1][1]:.4e}" if val is not None and net.validation else ""
            log(f"step {k:6d}/{steps} loss {L:.4e}{vtxt} |g| {gn:.3e} lr {cur:.2e} ({time.time() - t0:.1f} s)")
    if best[1] is not None:
        # the validation-based stop: the parameters of the best held-out loss
        net.set_params(best[1])
        net.best_step = best[2]
        if log:
            log(f"kept the parameters of step {best[2]} (validation loss {best[0]:.4e})")
    if val is not None and val[2] is not None:
        val[2].close()
    for _, tq in pool:
        tq.close()
    if tl is not None and not pool:
        tl.close()
    return net, hist


def main():
    ap = argparse.ArgumentParser(description="train the FNO preconditioner on one GPU")
    ap.add_argument("--N", type=int, default=256)
    ap.add_argument("--samples", type=int, default=32000,
                    help="32,000 at N = 256 serves N = 512 best on held-out systems; longer training fits "
                         "N = 256 at the expense of N = 512")
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1e-3)
    ap.add_argument("--seed", type=int, default=0)
    ap.add_argument("--identity-weight", type=float, default=0.0)
    ap.add_argument("--chain", type=int, default=3)
    ap.add_argument("--tiles", type=int, default=0,
                    help="train the network as the second level after T x T tile inverses (0: alone)")
    ap.add_argument("--map-pool", type=int, default=8,
                    help="with --tiles: batches of maps whose tile inverses are factored once and reused")
    ap.add_argument("--val-every", type=int, default=50,
                    help="evaluate a held-out set every this many steps and keep the best parameters (0: off)")
    ap.add_argument("--out", default="fno_params.npz")
    a = ap.parse_args()
    from .solver import Context
    ctx = Context(a.N)
    net, hist = train(ctx, a.samples, a.batch, a.lr, a.seed, log=lambda s: print(s, flush=True),
                      identity_weight=a.identity_weight, chain=a.chain, tiles_T=a.tiles,
                      map_pool=a.map_pool, val_every=a.val_every)
    fno.save(a.out, net.params.cpu().numpy(), N=a.N, samples=a.samples, seed=a.seed,
             loss=np.array([h[1] for h in hist], np.float32))
    print(f"saved {a.out}")


if __name__ == "__main__":
    main()
