"""Benchmark of the GMRES solve on one MI355X: solves/s to a true relative
residual of 1e-6 at N = 512, learned-preconditioned FGMRES against plain GMRES.

Workload: the plain-GMRES workload at N = 512 (262,144 unknowns, 1024 x 1024
FFT embedding), complex64, h = lambda/20, a seeded binary permittivity map of
8 x 8-cell blocks with values {1, 12.1} over the central 75 % of the area and
a plane wave at a seeded angle (cugmres_amd.problems.config_plain); x0 = 0;
a solve stops when the true ||b - A x|| <= 1e-6 ||b||.  All inputs are copied
to HBM before the timed region.  One step solves one seeded system; step k
uses seed --seed-offset + k (warm-up steps first).

Methods (each system is solved from scratch; the time of a solve includes
everything that depends on its map):
  - learned: flexible GMRES with the two-level preconditioner - the exact
    inverses of the operator on 32 x 32 tiles (--tile; factored per map on the
    device), then the FNO network as a correction on the residual they leave
    (cugmres_amd.tiles.TwoLevel) - with the long-recurrence path (the whole
    Krylov basis in HBM, up to --lr-restart steps).  The network is trained
    on the GPU before the timed region (or loaded with --fno-params).
  - tiles: the same solver with the tile inverses alone (the method without
    the network).
  - plain: GMRES(30) without a preconditioner, the reference's algorithm,
    capped at --max-cycles restart cycles.  On this physics it stalls near
    1e-2 and reaches 1e-6 on none of the systems; it is reported as measured.
`value` counts only solves that reached 1e-6 (0.0 if none did).

Printed: one JSON line (under 4 KB) with the metric, the three methods'
summaries, the roofline of the learned solve's dominant kernel and of the FFT
matvec (HIP events on the solver's stream), and the CPU baseline (the
oracle's complex64 GMRES(30), the C restatement of the reference's MKL path,
timed on the host; one core and all cores).  --details FILE writes the
per-system figures.
"""
import argparse
import ctypes
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "solves/s to 1e-6 rel. residual, N=512: learned-FGMRES vs plain GMRES"


def note(msg):
    """Progress on stderr (stdout carries only the result line)."""
    print(f"[bench {time.strftime('%H:%M:%S')}] {msg}", file=sys.stderr, flush=True)
HBM_PEAK_GBS = 8000.0


def parse(argv=None):
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=4)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--N", type=int, default=512)
    p.add_argument("--restart", type=int, default=30, help="restart of plain GMRES")
    p.add_argument("--tol", type=float, default=1e-6)
    p.add_argument("--max-cycles", type=int, default=100, help="cap of plain GMRES(restart)")
    p.add_argument("--tile", type=int, default=128, help="tile side of the tile inverses (8, 16, 32, 64, 128)")
    p.add_argument("--lr-restart", type=int, default=8192,
                   help="cycle length of the preconditioned solves (long-recurrence path)")
    p.add_argument("--lr-cycles", type=int, default=2)
    p.add_argument("--compare", type=int, default=2,
                   help="timed systems solved with the comparison methods (tiles, plain)")
    p.add_argument("--cpu-seconds", type=float, default=15.0, help="budget of the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--fno-params", default=None, help="trained network (.npz from cugmres_amd.learn)")
    p.add_argument("--train-samples", type=int, default=4000)
    p.add_argument("--train-N", type=int, default=512)
    p.add_argument("--seed-offset", type=int, default=0)
    p.add_argument("--details", default=None, help="write the per-system figures to this JSON file")
    p.add_argument("--batch", type=int, default=16, help="vectors of the batched / fno-step modes")
    p.add_argument("--instances", type=int, default=7,
                   help="c3 mode: instances (7 = one per grid size; 256 = the full batched workload)")
    p.add_argument("--mode", choices=["bench", "step", "copy", "batched", "fno-step", "copybw", "sweep", "cycles",
                                      "lr-step", "tiles-step", "c3", "c2", "fftref"],
                   default="bench",
                   help="bench: the benchmark; step: the per-kernel profile of one Arnoldi step of the fused "
                        "GMRES(--restart) kernels (for rocprofv3 --pmc); lr-step: the long-recurrence step kernels "
                        "at basis size --lr-j; tiles-step: 20 applications of the tile inverses; c3: the batched "
                        "workload (--instances maps of mixed N, 16 plane waves each); c2: the mixed-precision "
                        "workload (smooth map, FP64 refinement) at --N; copy: 3 copies of "
                        "256 MiB (counter calibration); batched: lockstep batches of 16 right-hand sides vs single "
                        "solves; fno-step: the per-kernel profile of one network forward pass on --batch vectors; "
                        "copybw: device-copy bandwidth at 8-256 MB; sweep: the reference's own benchmark (its "
                        "cylinder problem at N = 256..2048, Krylov dimension 3/30/49) through "
                        "cgm_Fast_GMRES_with_CUDA beside its published V100 / GTX 1080 Ti times; cycles: "
                        "3 x (2 restart cycles of cgm_fgmres_batched with a seeded untrained network, then 2 of "
                        "cgm_gmres_batched) on --batch systems")
    p.add_argument("--sweep-reps", type=int, default=20)
    p.add_argument("--lr-j", type=int, default=4096, help="lr-step mode: basis size of the profiled kernels")
    p.add_argument("--c2-seconds", type=float, default=600.0, help="c2 mode: no new outer step after this time")
    p.add_argument("--c3-method", choices=["block", "lockstep"], default="block",
                   help="c3 mode: block GMRES over each map's right-hand sides (one shared block Krylov space), "
                        "or flexible GMRES of every right-hand side in lockstep")
    p.add_argument("--block-restart", type=int, default=256, help="c3 block method: block steps per cycle")
    p.add_argument("--block-cycles", type=int, default=16, help="c3 block method: cap on the cycles")
    p.add_argument("--tiles-synthetic", action="store_true",
                   help="tiles-step mode: time the application on synthetic inverses (no factorisation)")
    return p.parse_args(argv)


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus %d is not supported "
                                   "(a solve does not shard across GPUs; batches of systems would shard "
                                   "by system with no communication)" % a.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    from cugmres_amd import problems, solver

    if a.mode == "copy":
        src = torch.ones(32 * 1024 * 1024, dtype=torch.complex64, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes": src.numel() * 8, "copies": 3}))
        return
    if a.mode == "batched":
        print(json.dumps(batched_throughput(a.restart)), flush=True)
        return
    if a.mode == "sweep":
        from cugmres_amd import sweep
        r = sweep.run(reps=a.sweep_reps, log=lambda s: print(s, flush=True))
        print(json.dumps({"sweep": {f"{N}x{m}": v for (N, m), v in r.items()}}), flush=True)
        return
    if a.mode == "fftref":
        print(json.dumps(fft_reference(a), separators=(",", ":")), flush=True)
        return
    if a.mode == "c3":
        print(json.dumps(c3_run(a), separators=(",", ":")), flush=True)
        return
    if a.mode == "c2":
        print(json.dumps(c2_run(a), separators=(",", ":")), flush=True)
        return
    if a.mode == "copybw":
        out = {}
        for mb in (8, 16, 32, 64, 256):
            src = torch.ones(mb * 1024 * 1024 // 8, dtype=torch.complex64, device="cuda")
            dst = torch.empty_like(src)
            for _ in range(3):
                dst.copy_(src)
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record()
            for _ in range(20):
                dst.copy_(src)
            e1.record()
            torch.cuda.synchronize()
            us = e0.elapsed_time(e1) / 20 * 1e3
            out[f"{mb}MiB"] = {"us": us, "GBps_read_plus_write": 2 * src.numel() * 8 / (us * 1e-6) / 1e9}
        print(json.dumps({"copy_bandwidth": out}), flush=True)
        return
    N = a.N
    ctx = solver.Context(N, 0)
    if a.mode == "fno-step":
        from cugmres_amd import fno
        B = a.batch
        cb = [problems.config_plain(N, seed=s_)[:2] for s_ in range(B)]
        net = fno.Preconditioner(ctx, batch=B, params=fno.init_params(0, zero_out=False))
        dchi = ctx.to_device(np.stack([c.reshape(-1) for c, _ in cb]), "f32")
        db = ctx.to_device(np.stack([b.reshape(-1) for _, b in cb]))
        print(json.dumps(profile_fno(ctx, net, dchi, db, B)))
        return
    if a.mode == "lr-step":
        print(json.dumps(profile_longrec(ctx, a.lr_j)), flush=True)
        return
    if a.mode == "tiles-step":
        from cugmres_amd import tiles
        if a.tiles_synthetic:
            # the application alone, on synthetic inverses (no rocSOLVER call)
            from cugmres_amd._lib import lib
            tl = tiles.Tiles(ctx, a.tile, 1)
            ms = ctypes.c_float()
            lib().cgm_profile_tiles(tl.handle, 20, ctypes.byref(ms))
            print(json.dumps({"tile": a.tile, "us": ms.value * 1e3}), flush=True)
            return
        chi, b, _ = problems.config_plain(N, seed=0)
        tl = tiles.Tiles(ctx, a.tile, 1).factor(ctx.block_helmholtz(problems.KH), ctx.to_device(chi.reshape(-1), "f32"))
        print(json.dumps(profile_tiles(ctx, tl, ctx.to_device(b.reshape(-1)))), flush=True)
        return
    if a.mode == "cycles":
        from cugmres_amd import fno
        B = a.batch
        g = ctx.spectrum_helmholtz(problems.KH)
        cb = [problems.config_plain(N, seed=s_)[:2] for s_ in range(B)]
        dchi = ctx.to_device(np.stack([c.reshape(-1) for c, _ in cb]), "f32").reshape(B, N * N)
        db = ctx.to_device(np.stack([b.reshape(-1) for _, b in cb])).reshape(B, N * N)
        p = fno.init_params(3, zero_out=False)
        fno.unpack(p)["W_out"][...] *= 0.05
        net = fno.Preconditioner(ctx, batch=B, params=p)
        out = {"learned_ms": [], "plain_ms": []}
        for _ in range(3):
            t0 = time.perf_counter()
            ctx.fgmres_batched(g, dchi, db, precond=net, restart=a.restart, tol=1e-9, max_cycles=2)
            torch.cuda.synchronize()
            t1 = time.perf_counter()
            ctx.gmres_batched(g, dchi, db, restart=a.restart, tol=1e-9, max_cycles=2)
            torch.cuda.synchronize()
            out["learned_ms"].append((t1 - t0) * 1e3)
            out["plain_ms"].append((time.perf_counter() - t1) * 1e3)
        print(json.dumps({"cycles": out, "batch": B, "N": N, "restart": a.restart}), flush=True)
        return
    if a.mode == "step":
        chi, b, _ = problems.config_plain(N, seed=0)
        prof = profile_kernels(ctx, ctx.spectrum_helmholtz(problems.KH),
                               ctx.to_device(chi.reshape(-1), "f32"), ctx.to_device(b.reshape(-1)),
                               a.restart)
        print(json.dumps(prof))
        return
    out, details = run_bench(a, ctx)
    if a.details:
        with open(a.details, "w") as f:
            json.dump(details, f)
    line = json.dumps(out, separators=(",", ":"))
    print(line, flush=True)


def _stats(vals):
    import numpy as np
    v = np.asarray([x for x in vals if x is not None], float)
    if v.size == 0:
        return None
    return [float(v.min()), float(np.median(v)), float(v.max())]


def summarize_method(infos, wall_s, name):
    """Compact summary of one method's timed solves (no per-system arrays):
    converged count, solves/s counting converged solves only, and
    [min, median, max] of iterations, residuals and per-solve times."""
    conv = int(sum(1 for i in infos if i["converged"]))
    return {"method": name, "converged": conv, "of": len(infos),
            "solves_per_s": conv / wall_s if wall_s > 0 else 0.0,
            "wall_s": wall_s,
            "iterations": _stats([i["iterations"] for i in infos]),
            "rel_residual": _stats([i["rel_residual"] for i in infos]),
            "ms_per_solve": _stats([i["ms"] for i in infos])}


def run_bench(a, ctx):
    import numpy as np
    import torch
    from cugmres_amd import fno, learn, problems, solver, tiles

    N = ctx.N
    g, g64 = ctx.spectrum_helmholtz(problems.KH, fp64=True)
    ctx.set_flexible_fp64(g64)   # FP64 products A z_j of the preconditioned vectors
    blk = ctx.block_helmholtz(problems.KH)
    total = a.warmup + a.steps
    inputs = []
    for k in range(total):
        chi, b, _ = problems.config_plain(N, seed=a.seed_offset + k)
        inputs.append((ctx.to_device(chi.reshape(-1), "f32"), ctx.to_device(b.reshape(-1))))
    x = torch.zeros(N * N, dtype=torch.complex64, device=ctx.device)

    # ---- the network (trained before the timed region) ---------------------
    if a.fno_params:
        params = fno.load(a.fno_params)
        training = {"loaded": os.path.basename(a.fno_params)}
    else:
        tctx = solver.Context(a.train_N, 0)
        t0 = time.perf_counter()
        note(f"training the network: {a.train_samples} samples at N = {a.train_N}")
        # pooled maps: 4 batches of 8 (32 maps), each batch's tile inverses factored once (20 GB of
        # W per batch at N = 512, T = 128); a short training at the solve's own N: longer ones lower
        # the loss but raise the FGMRES step count
        tnet, hist = learn.train(tctx, samples=a.train_samples, batch=8, lr=1e-3, seed=0, tiles_T=a.tile,
                                 map_pool=4, log=note, log_every=200, val_every=50)
        torch.cuda.synchronize()
        note(f"trained in {time.perf_counter() - t0:.1f} s")
        params = tnet.params.cpu().numpy()
        training = {"N": a.train_N, "samples": a.train_samples, "tiles": a.tile,
                    "seconds": round(time.perf_counter() - t0, 1), "loss_first": hist[0][1],
                    "loss_last100": float(np.mean([h[1] for h in hist[-100:]])),
                    "validation_best": min(c[1] for c in tnet.validation), "kept_step": tnet.best_step,
                    "validation_last": tnet.validation[-1][1]}
        tnet.close()
        tctx.close()
    tl = tiles.Tiles(ctx, a.tile, 1)
    net = fno.Preconditioner(ctx, batch=1, params=params)
    two = tiles.TwoLevel(tl, net)
    torch.cuda.synchronize()

    def solve(method, k):
        dchi, db = inputs[k]
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        if method == "plain":
            info = ctx.gmres(g, dchi, db, x.zero_(), restart=a.restart, tol=a.tol, max_cycles=a.max_cycles)[1]
        else:
            tl.factor(blk, dchi)   # the inverses depend on the map: part of the solve
            info = ctx.fgmres(g, dchi, db, x.zero_(), precond=two if method == "learned" else tl,
                              restart=a.lr_restart, tol=a.tol, max_cycles=a.lr_cycles, stall_ratio=0.0)[1]
        torch.cuda.synchronize()
        info["ms"] = (time.perf_counter() - t0) * 1e3
        info["seed"] = a.seed_offset + k
        note(f"{method} seed {info['seed']}: {info['iterations']} steps, converged {info['converged']}, "
             f"residual {info['rel_residual']:.2e}, {info['ms']:.0f} ms")
        return info

    def run(method, warm, timed, live=False):
        for k in warm:
            solve(method, k)
        torch.cuda.synchronize()
        if live:
            ctx.longrec_timing(True)
        t0 = time.perf_counter()
        infos = [solve(method, k) for k in timed]
        wall = time.perf_counter() - t0
        if live:
            live_t = ctx.longrec_timing_read()
            ctx.longrec_timing(False)
            return infos, wall, live_t
        return infos, wall

    timed = list(range(a.warmup, total))
    cmp_sys = timed[:max(0, min(a.compare, len(timed)))]
    l_infos, l_wall, live = run("learned", range(a.warmup), timed, live=True)
    t_infos, t_wall = run("tiles", range(min(1, a.warmup)), cmp_sys) if cmp_sys else ([], 0.0)
    p_infos, p_wall = run("plain", range(min(1, a.warmup)), cmp_sys) if cmp_sys else ([], 0.0)
    learned = summarize_method(l_infos, l_wall, "learned")
    tiles_s = summarize_method(t_infos, t_wall, "tiles") if t_infos else None
    plain = summarize_method(p_infos, p_wall, "plain") if p_infos else None

    # ---- rooflines: the learned solve's kernels at its mean basis size ----------
    iters = learned["iterations"][1] if learned["iterations"] else 1000
    jmean = max(16, int(iters // 2))
    lr = profile_longrec(ctx, jmean)
    dchi0, db0 = inputs[0]
    tl.factor(blk, dchi0)
    tp = profile_tiles(ctx, tl, db0)
    mv = profile_kernels(ctx, g, dchi0, db0, a.restart)["kernels"]["matvec_total"]
    fp = profile_fno(ctx, net, dchi0.reshape(1, -1), db0.reshape(1, -1), 1)["kernels"]["fno_forward_total"]
    per_step = {"lr_proj": lr["lr_proj"]["us"], "lr_update": lr["lr_update"]["us"],
                "lr_reduce_givens": lr["lr_reduce"]["us"] + lr["lr_givens"]["us"],
                "tile_apply": tp["us"], "matvec_x2": 2 * mv["us"], "fno_forward": fp["us"]}
    roof = live_roofline(live, N, per_step, jmean, a.tile)
    matvec = {"us": mv["us"], "bytes": mv["bytes"], "GBps": mv["GBps"], "frac": mv["GBps"] / HBM_PEAK_GBS,
              "per_unit": "100 N^2 B per matvec (quarter spectrum)",
              "batched16": batched_matvec_rate(ctx, g, dchi0)}

    note("profiles done; reference problem")
    refp = reference_problem(ctx, a.restart) if N == 512 else None
    note("CPU baseline")
    cpu = None if a.no_cpu else cpu_baseline(N, a.restart, a.cpu_seconds, a.max_cycles)

    out = make_line(a, N, ctx.M, learned, tiles_s, plain, l_wall, training, roof, matvec, cpu, refp)
    details = {"learned": l_infos, "tiles": t_infos, "plain": p_infos, "profiles": {"longrec": lr, "tiles": tp}}
    for v in details.values():
        if isinstance(v, list):
            for i in v:
                i.pop("history", None)
    return out, details


def make_line(a, N, M, learned, tiles_s, plain, l_wall, training, roof, matvec, cpu, refp):
    """The one JSON object bench.py prints (summaries only, no per-system
    arrays, so the line stays far below 4 KB at any --steps)."""
    return {
        "metric": METRIC, "value": learned["solves_per_s"], "unit": "solves/s", "n_gpus": 1,
        "steps": a.steps, "warmup": a.warmup, "ms_per_step": l_wall / max(1, a.steps) * 1e3,
        "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "c64",
        "data": "synthetic (seeded eps maps and plane waves)",
        "config": {"workload": f"N={N} 8x8-block eps in {{1,12.1}} over 75% area, h=lambda/20, tol {a.tol:g}, "
                               f"one seeded system per step; learned = FGMRES({a.lr_restart}) with {a.tile}x{a.tile} "
                               f"tile inverses + FNO; plain = GMRES({a.restart}) capped at {a.max_cycles} cycles",
                   "N": N, "unknowns": N * N, "embedding": M, "tile": a.tile, "lr_restart": a.lr_restart,
                   "plain_restart": a.restart, "tol": a.tol},
        "learned": learned, "tiles": tiles_s, "plain": plain,
        "speedup_vs_plain": (learned["solves_per_s"] / plain["solves_per_s"])
        if plain and plain["solves_per_s"] > 0 else None,
        "speedup_vs_tiles": (learned["solves_per_s"] / tiles_s["solves_per_s"])
        if tiles_s and tiles_s["solves_per_s"] > 0 else None,
        "training": training, "roofline": roof, "matvec": matvec, "cpu_baseline": cpu, "reference_problem": refp,
    }


C3_SIZES = (64, 96, 128, 192, 256, 384, 512)


def c3_instances(count, seed=2024):
    """The batched workload's instances: grid size uniform in C3_SIZES (in
    order when count <= 7, one per size), block size uniform in 4..8, a
    seeded binary map (p = 0.5, central 75 %) per instance."""
    from cugmres_amd import problems
    rng = problems.SplitMix64(seed)
    out = []
    for i in range(count):
        N = C3_SIZES[i % 7] if count <= 7 else C3_SIZES[min(6, int(rng.uniform() * 7))]
        blk = 4 + min(4, int(rng.uniform() * 5))
        out.append((N, blk, seed * 1000 + i))
    return out


def batched_matvec_rate(ctx, g, chi, B=16, reps=20):
    """cgm_matvec_batched on B vectors of one map with the autotuner's radix
    plan, timed with HIP events on the context's stream: microseconds per
    launch and the fraction of HBM peak on 88 N^2 B per vector plus the map
    and the quarter spectrum once (the bytes fft_reference counts)."""
    import torch
    N, M = ctx.N, ctx.M
    x = torch.randn(B, N * N, dtype=torch.complex64, device=ctx.device)
    y = torch.empty_like(x)
    for _ in range(3):
        ctx.matvec_batched(g, chi, x, 0, y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        ctx.matvec_batched(g, chi, x, 0, y)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / reps * 1e3
    nbytes = B * 88 * N * N + 4 * N * N + 8 * (M // 2 + 1) ** 2
    return {"us": round(us, 1), "frac": round(nbytes / (us * 1e-6) / 1e9 / HBM_PEAK_GBS, 3)}


def fft_reference(a, B=16, reps=20):
    """The library's batched matvec (three fused FFT passes, B vectors of one
    map at --N) beside rocFFT (torch.fft, hipFFT) doing the transforms alone:
    a forward and an inverse 2-D FFT of B (2N x 2N) complex64 arrays, and the
    same on the N x 2N row-transformed half.  Bytes are the minimum each
    moves: the matvec 88 N^2 per vector (x in twice, the N x 2N intermediate
    written and read twice, y out) plus the map and the quarter spectrum once;
    a 2-D FFT of an M x M array reads and writes it once per pass (rocFFT's
    two passes: 32 M^2 B)."""
    import torch
    from cugmres_amd import problems, solver
    N = a.N
    ctx = solver.Context(N, 0)
    M = ctx.M
    g = ctx.spectrum_helmholtz(problems.KH)
    chi = ctx.to_device(problems.config_plain(N, seed=0)[0].reshape(-1), "f32")
    x = torch.randn(B, N * N, dtype=torch.complex64, device=ctx.device)
    y = torch.empty_like(x)

    def timed(fn):
        for _ in range(3):
            fn()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(reps):
            fn()
        e1.record()
        torch.cuda.synchronize()
        return e0.elapsed_time(e1) / reps * 1e3

    out = {"N": N, "M": M, "B": B, "radix_plan": ctx.radix_plan()}
    us = timed(lambda: ctx.matvec_batched(g, chi, x, 0, y))
    mv_bytes = B * 88 * N * N + 4 * N * N + 8 * (M // 2 + 1) ** 2
    out["matvec_batched"] = {"us": us, "bytes": mv_bytes, "GBps": mv_bytes / (us * 1e-6) / 1e9}
    if M == 1024:   # both pass-2 kernels, whatever the autotuner chose
        auto = out["radix_plan"]["wide_from_batch"]
        for name, wide in (("lds_radix8", 0), ("one_wave", 1)):
            ctx.set_radix_plan(wide)
            u = timed(lambda: ctx.matvec_batched(g, chi, x, 0, y))
            out["matvec_batched_" + name] = {"us": u, "GBps": mv_bytes / (u * 1e-6) / 1e9}
        ctx.set_radix_plan(auto)
    big = torch.randn(B, M, M, dtype=torch.complex64, device=ctx.device)
    us = timed(lambda: torch.fft.ifft2(torch.fft.fft2(big)))
    b2 = 2 * 32 * M * M * B
    out["rocfft_fft2_ifft2_MxM"] = {"us": us, "bytes": b2, "GBps": b2 / (us * 1e-6) / 1e9}
    us = timed(lambda: torch.fft.fft2(big))
    out["rocfft_fft2_MxM"] = {"us": us, "bytes": b2 // 2, "GBps": b2 / 2 / (us * 1e-6) / 1e9}
    ctx.close()
    return out


def c3_run(a, rhs=16, mem_budget=150e9):
    """Batched workload: per instance one map and `rhs` plane waves at angles
    2 pi k / rhs.  --c3-method block (default): block GMRES over the map's
    right-hand sides with its tile inverses (cugmres_amd.block.solve_map: one
    shared block Krylov space, the projection V^H W as batched complex GEMMs,
    cycles of --block-restart block steps that restart from the FP64 true
    residual with the solution kept in complex128).  --c3-method lockstep: the
    right-hand sides are solved in lockstep by flexible GMRES with the tile inverses
    of that map (factored once; the dense form streams each inverse once per
    four right-hand sides, the reduced form runs the systems of one W block
    back to back on one XCD) on the long-recurrence path with cycles of --lr-restart
    steps (two at most).  The right-hand sides run in lockstep groups as large
    as mem_budget bytes of Krylov bases allow (all 16 up to N = 256, fewer
    above).  A right-hand side left above the tolerance (on the hardest maps
    the complex64 rounding of the solution alone leaves about 6e-8 ||A|| ||x||)
    is then refined in FP64 from where it stopped (cgm_gmres_refine_pc with
    the same tiles, corrections to 1e-1), inside the timed region.  Also
    plain GMRES(30) on the first instance of each size (capped at 100
    cycles).  Reports RHS-solves/s counting only right-hand sides that reached
    the tolerance (the FP64 true residual for the refined ones)."""
    import numpy as np
    import torch
    from cugmres_amd import block, problems, solver, tiles
    ctxs = {}
    per_n = {}
    t_all = 0.0
    conv_all = 0
    total = 0
    # grouped by N, one context at a time: each context keeps its Krylov
    # workspace (up to mem_budget bytes) until it is closed
    for (N, blk_cells, seed) in sorted(c3_instances(a.instances), key=lambda t: t[0]):
        if N not in ctxs:
            for c_old, _, _, _ in ctxs.values():
                c_old.close()
            ctxs.clear()
            c = solver.Context(N, 0)
            g32, g64 = c.spectrum_helmholtz(problems.KH, fp64=True)
            c.set_flexible_fp64(g64)
            ctxs[N] = (c, g32, c.block_helmholtz(problems.KH), g64)
        ctx, g, gb, _ = ctxs[N]
        eps = problems.block_eps_map(N, blk_cells, 0.5, 0.75, seed=seed)
        chi = ctx.to_device((eps - 1.0).reshape(-1), "f32")
        b = torch.stack([ctx.to_device(problems.plane_wave(N, problems.KH, 2 * math.pi * k / rhs).reshape(-1))
                         for k in range(rhs)])
        m = a.lr_restart
        per_rhs = 2 * 8 * N * N * (m + 1) + 16 * m * m   # basis, preconditioned vectors, R
        grp = int(max(1, min(rhs, mem_budget // per_rhs)))
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        T = max(t for t in (a.tile, 64, 32, 16) if t <= a.tile and N % t == 0)   # largest tile that divides N
        tl = tiles.Tiles(ctx, T, 1).factor(gb, chi)
        infos = []
        refined = 0
        bstat = None
        if a.c3_method == "block":
            chi64 = ctx.to_device((eps - 1.0).reshape(-1), "f64")
            X, binfo = block.solve_map(ctx, g, chi, b, tiles=tl, restart=a.block_restart, tol=a.tol,
                                       max_cycles=a.block_cycles, gamma64=ctxs[N][3], chi64=chi64)
            infos = [{"converged": bool(cv), "rel_residual": float(rr), "iterations": binfo["block_steps"]}
                     for cv, rr in zip(binfo["converged"], binfo["rel_residual"])]
            bstat = {k: binfo[k] for k in ("block_steps", "cycles", "breakdowns")}
            grp = rhs
        for g0 in range(0, rhs if a.c3_method == "lockstep" else 0, grp):
            bg = b[g0:g0 + grp].contiguous()
            xg, ig = ctx.fgmres_batched(g, chi, bg, precond=tl, restart=m, tol=a.tol, max_cycles=2)
            # a system left above the tolerance at the complex64 floor of its solution (about 6e-8
            # ||A|| ||x|| per entry) is refined in FP64 from where it stopped, with the same tiles
            for k, inf in enumerate(ig):
                if inf["converged"]:
                    continue
                chi64 = ctx.to_device((eps - 1.0).reshape(-1), "f64")
                b64 = bg[k].to(torch.complex128)
                x64 = xg[k].to(torch.complex128).contiguous()
                _, rinf = ctx.gmres_refine_pc(g, ctxs[N][3], chi64, b64, tl, x=x64, inner_restart=m,
                                              inner_tol=1e-1, inner_max_cycles=1, outer_tol=a.tol, max_outer=3)
                inf["converged"] = rinf["converged"]
                inf["rel_residual"] = rinf["rel_residual"]
                inf["iterations"] += rinf["inner_iterations"]
                refined += 1
            infos += ig
            if N >= 384:
                note(f"c3 N={N}: right-hand sides {g0}-{g0 + len(ig) - 1} done, "
                     f"{sum(1 for i in ig if i['converged'])} converged, {time.perf_counter() - t0:.1f} s")
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        tl.close()
        conv = sum(1 for i in infos if i["converged"])
        e = per_n.setdefault(N, {"instances": 0, "rhs": 0, "converged": 0, "seconds": 0.0, "restart": m,
                                 "lockstep_group": grp, "tile": T, "refined_fp64": 0, "iterations": [],
                                 "residuals": []})
        e["refined_fp64"] += refined
        if bstat is not None:
            e.setdefault("block", []).append(bstat)
        e["instances"] += 1
        e["rhs"] += rhs
        e["converged"] += conv
        e["seconds"] += dt
        e["iterations"] += [i["iterations"] for i in infos]
        e["residuals"] += [i["rel_residual"] for i in infos]
        if e["instances"] == 1:
            t1 = time.perf_counter()
            _, pinfos = ctx.gmres_batched(g, chi, b, restart=30, tol=a.tol, max_cycles=100)
            torch.cuda.synchronize()
            e["plain_gmres30"] = {"converged": sum(1 for i in pinfos if i["converged"]), "of": rhs,
                                  "seconds": time.perf_counter() - t1,
                                  "rel_residual": _stats([i["rel_residual"] for i in pinfos])}
        t_all += dt
        conv_all += conv
        total += rhs
        note(f"c3 N={N} block={blk_cells} converged {conv}/{rhs} in {dt:.2f} s ({a.c3_method}; "
             + (f"{bstat['block_steps']} block steps in {bstat['cycles']} cycles)" if bstat else
                f"{refined} refined in FP64; groups of {grp}; iterations "
                f"{_stats([i['iterations'] for i in infos])})"))
    for e in per_n.values():
        e["iterations"] = _stats(e["iterations"])
        e["residuals"] = _stats(e["residuals"])
    for c, _, _, _ in ctxs.values():
        c.close()
    return {"workload": f"{a.instances} instances x {rhs} plane waves, N in {list(C3_SIZES)}, blocks 4-8, "
                        f"eps in {{1,12.1}}, tol {a.tol:g}",
            "method": ("block GMRES over each map's 16 right-hand sides with its tile inverses (one "
                       "factorisation per map), projection V^H W as batched complex GEMMs, cycles restarting from "
                       "the FP64 true residual" if a.c3_method == "block" else
                       "FGMRES with tile inverses, lockstep, one factorisation per map; FP64 refinement from "
                       "where it stopped for a system left at the complex64 floor"),
            "rhs_solves_per_s": conv_all / t_all if t_all else 0.0,
            "converged": conv_all, "of": total, "seconds": t_all, "per_N": per_n}


def c2_run(a):
    """Mixed-precision workload: smooth map (sigma = 4 cells, eps in [1, 12.1])
    and plane wave at --N; FP64 refinement whose corrections are solved by
    complex64 flexible GMRES with the T x T tile inverses (long-recurrence
    path).  Reports the FP64 true relative residual per outer step and the
    time; the error against an FP64 solve is checked by the tests at sizes the
    CPU oracle covers."""
    import torch
    from cugmres_amd import problems, solver, tiles
    N = a.N
    ctx = solver.Context(N, 0)
    g32, g64 = ctx.spectrum_helmholtz(problems.KH, fp64=True)
    ctx.set_flexible_fp64(g64)
    blk = ctx.block_helmholtz(problems.KH)
    out = {"workload": f"N={N} smooth eps in [1,12.1] (sigma 4 cells), plane wave, FP64 refinement to 1e-10, "
                       f"inner FGMRES with {a.tile}x{a.tile} tile inverses to 1e-4", "solves": []}
    for s in range(a.steps):
        chi, b, _ = problems.config_mixed(N, seed=a.seed_offset + s)
        chi32 = ctx.to_device(chi.reshape(-1), "f32")
        chi64 = ctx.to_device(chi.reshape(-1), "f64")
        b64 = ctx.to_device(b.reshape(-1), "c128")
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        tl = tiles.Tiles(ctx, a.tile, 1).factor(blk, chi32)
        # one correction per call, so that each outer step reports itself;
        # no new step starts once --c2-seconds have passed
        x = torch.zeros_like(b64)
        info = {"outer": 0, "inner_iterations": 0, "converged": False, "rel_residual": 1.0, "steps": []}
        for _ in range(12):
            x, st = ctx.gmres_refine_pc(g32, g64, chi64, b64, tl, x=x, inner_restart=a.lr_restart, inner_tol=1e-4,
                                        inner_max_cycles=1, outer_tol=1e-10, max_outer=1)
            info["outer"] += st["outer"]
            info["inner_iterations"] += st["inner_iterations"]
            info["converged"] = st["converged"]
            info["rel_residual"] = st["rel_residual"]
            info["steps"].append([st["inner_iterations"], st["rel_residual"]])
            note(f"c2 N={N} seed {a.seed_offset + s}: outer step {info['outer']}, inner {st['inner_iterations']} "
                 f"steps, FP64 residual {st['rel_residual']:.2e}, {time.perf_counter() - t0:.1f} s")
            if st["converged"] or st["outer"] == 0 or time.perf_counter() - t0 > a.c2_seconds:
                break
        torch.cuda.synchronize()
        dt = time.perf_counter() - t0
        # FP64 true residual of the result, recomputed independently
        r = b64 - ctx.matvec_c128(g64, chi64, x)
        rel_fp64 = float(torch.linalg.vector_norm(r) / torch.linalg.vector_norm(b64))
        tl.close()
        out["solves"].append({"seed": a.seed_offset + s, "seconds": dt, "fp64_rel_residual": rel_fp64, **info})
        note(f"c2 N={N} seed {a.seed_offset + s}: {info} fp64 residual {rel_fp64:.2e} {dt:.2f} s")
    ctx.close()
    return out


def profile_longrec(ctx, j, reps=20):
    """Durations of the long-recurrence step kernels at basis size j (HIP
    events, cgm_profile_longrec) with their algorithmic bytes: the projection
    reads j basis vectors and u, w once (8 B each per cell); the update reads
    j basis vectors, u and w and writes q_j and u_{j+1}."""
    from cugmres_amd import _lib
    from cugmres_amd._lib import lib
    ms = (ctypes.c_float * 4)()
    _lib.check(lib().cgm_profile_longrec(ctx.handle, int(j), int(reps), ms), "cgm_profile_longrec")
    n = ctx.N * ctx.N
    b_proj = 8 * n * (j + 2)
    b_upd = 8 * n * (j + 4)
    us = [ms[i] * 1e3 for i in range(4)]
    mk = lambda k, u, b: {"kernel": k, "us": u, "bytes": b, "GBps": b / (u * 1e-6) / 1e9 if b else None,
                          "per_unit": f"{b} B per launch: 8 B x N^2 x (j + {2 if k == 'k_lr_proj' else 4}), "
                                      f"j = {j}, N = {ctx.N}" if b else None}
    return {"j": j, "lr_proj": mk("k_lr_proj", us[0], b_proj), "lr_reduce": mk("k_lr_red", us[1], 0),
            "lr_givens": mk("k_lr_fin", us[2], 0), "lr_update": mk("k_lr_upd", us[3], b_upd)}


def profile_tiles(ctx, tl, v, reps=20):
    """Duration of one application of the tile inverses (torch events on the
    context's stream) with its algorithmic bytes (cgm_tiles_apply_bytes: the
    inverse columns streamed plus the vector in and out)."""
    import torch
    y = torch.empty_like(v)
    for _ in range(3):
        tl.apply(v, out=y)
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(reps):
        tl.apply(v, out=y)
    e1.record()
    torch.cuda.synchronize()
    us = e0.elapsed_time(e1) / reps * 1e3
    N, T = ctx.N, tl.T
    b = tl.apply_bytes(1)
    full = (N // T) ** 2 * T ** 4 * (4 if T == 64 else 8)
    return {"kernel": "tile application (k_tile_ws, k_tile_wsum, FFT convolution, combine)" if T >= 64
            else "k_tile_apply",
            "us": us, "bytes": b,
            "GBps": b / (us * 1e-6) / 1e9, "kept_fraction": (b - 16 * N * N) / full,
            "per_unit": f"{b:.0f} B per launch: kept inverse columns + 16 N^2 B, N = {N}, T = {T}"}


def batched_throughput(restart, cycles=3, B=16):
    """Batched workload shape: B plane waves at evenly spaced angles on one
    seeded binary map (8 x 8 blocks, eps {1, 12.1}, 75 % area) per grid size,
    `cycles` full GMRES(restart) cycles (tolerance 0, so every system runs
    every step).  Times one cgm_gmres_batched call against B cgm_gmres calls
    on the same systems."""
    import numpy as np
    import torch
    from cugmres_amd import problems, solver
    out = {"workload": f"{B} plane waves x one map, GMRES({restart}), {cycles} cycles", "sizes": {}}
    for N in (64, 96, 128, 192, 256, 384, 512):
        ctx = solver.Context(N, 0)
        g = ctx.spectrum_helmholtz(problems.KH)
        eps = problems.block_eps_map(N, 8, 0.5, 0.75, seed=N)
        chi = ctx.to_device((eps - 1.0).reshape(-1), "f32")
        b = torch.stack([ctx.to_device(problems.plane_wave(N, problems.KH, 2 * math.pi * k / B).reshape(-1))
                         for k in range(B)])
        x = torch.zeros_like(b)
        ctx.gmres_batched(g, chi, b, x, restart=restart, tol=0.0, max_cycles=1)
        torch.cuda.synchronize()
        t = time.perf_counter()
        _, infos = ctx.gmres_batched(g, chi, b, x.zero_(), restart=restart, tol=0.0, max_cycles=cycles)
        torch.cuda.synchronize()
        tb = time.perf_counter() - t
        steps = sum(i["iterations"] for i in infos)
        t = time.perf_counter()
        for k in range(B):
            ctx.gmres(g, chi, b[k], x[k].zero_(), restart=restart, tol=0.0, max_cycles=cycles)
        torch.cuda.synchronize()
        ts = time.perf_counter() - t
        # batched matvec: one launch per FFT pass for all B vectors
        y = torch.empty_like(b)
        for _ in range(3):
            ctx.matvec_batched(g, chi, b, 0, y)
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(20):
            ctx.matvec_batched(g, chi, b, 0, y)
        e1.record()
        torch.cuda.synchronize()
        us_mv = e0.elapsed_time(e1) / 20 * 1e3
        M = ctx.M
        # per vector: x twice (passes 1 and 3) and y once (8 B each), the N x M
        # scratch written and read by passes 1-2-3 (8 B x 4); once per batch:
        # the shared chi (4 B) and the quarter spectrum
        mv_bytes = B * (24 * N * N + 32 * N * M) + 4 * N * N + 8 * (M // 2 + 1) ** 2
        out["sizes"][N] = {"batched_ms": tb * 1e3, "single_ms": ts * 1e3, "speedup": ts / tb,
                           "system_steps": steps, "us_per_system_step_batched": tb / steps * 1e6,
                           "us_per_system_step_single": ts / steps * 1e6,
                           "batched_matvec_us": us_mv, "batched_matvec_GBps": mv_bytes / (us_mv * 1e-6) / 1e9,
                           "batched_matvec_hbm_frac": mv_bytes / (us_mv * 1e-6) / 1e9 / HBM_PEAK_GBS}
        ctx.close()
    return out


LIVE_KERNELS = {"proj": "k_lr_proj", "upd": "k_lr_upd", "tile": "k_tile_apply"}


def live_roofline(live, N, per_step, jmean, tile=64):
    """Roofline of the dominant kernel of the timed learned solves, from the
    live HIP-event sums (cgm_longrec_timing over the timed region, on the
    solver's stream): achieved = algorithmic bytes summed over every launch /
    summed launch durations.  traffic = the HBM bytes per launch measured by
    rocprofv3 (profiles/hbm_traffic_n512.json: the basis kernels at basis
    size 4096, scaled to this run's mean algorithmic bytes per launch; the
    tile application at this tile size)."""
    key = max((k for k in LIVE_KERNELS if k in live), key=lambda k: live[k]["ms"])
    t = live[key]
    n = N * N
    ach = t["bytes"] / (t["ms"] * 1e-3) / 1e9 if t["ms"] > 0 else 0.0
    per_launch = t["bytes"] / max(1, t["launches"])
    traffic = None
    if key == "tile":
        if tile >= 64:
            kern = "k_tile_ws"
            ref = measured_traffic(f"{kern}_T{tile}", N)   # synthetic full tiles: scale by the bytes
            if ref:
                traffic = ref["bytes_per_launch"] * per_launch / ref["algorithmic"]
            per_unit = (f"the stored lower triangle of every W_t, b(b+1)/2 blocks of 128 x 128 x 4 B (FP16, "
                        f"b = ceil(n_t / 128), n_t = cells with chi != 0 padded to 16) "
                        f"+ 16 B x N^2 per launch (T = {tile}, N = {N}): {per_launch:.4g} B, "
                        f"{t['launches']} launches of the timed solves")
        else:
            kern = "k_tile_apply"
            ref = measured_traffic(f"{kern}_T{tile}", N)
            traffic = ref["bytes_per_launch"] if ref else None
            per_unit = (f"(N/T)^2 T^4 x 8 B of inverses + 16 B x N^2 per launch (T = {tile}, N = {N}): "
                        f"{per_launch:.4g} B, {t['launches']} launches of the timed solves")
    else:
        kern = LIVE_KERNELS[key]
        extra = 2 if key == "proj" else 4
        jbar = per_launch / (8 * n) - extra
        ref = measured_traffic(kern + "_j4096", N)
        if ref:
            traffic = ref["bytes_per_launch"] * per_launch / (8 * n * (4096 + extra))
        per_unit = (f"8 B x N^2 x (j + {extra}) per launch at basis size j (N = {N}); mean j {jbar:.0f} "
                    f"over the {t['launches']} launches of the timed solves")
    return {"bound": "hbm", "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": ach / HBM_PEAK_GBS,
            "traffic": traffic, "kernel": kern, "launches": t["launches"],
            "us": t["ms"] * 1e3 / max(1, t["launches"]),
            "per_unit": per_unit,
            "other": {LIVE_KERNELS[k]: {"GBps": round(live[k]["bytes"] / (live[k]["ms"] * 1e-3) / 1e9)
                                        if live[k]["ms"] > 0 else None,
                                        "ms_total": round(live[k]["ms"], 1), "launches": live[k]["launches"]}
                      for k in LIVE_KERNELS if k != key and k in live},
            "us_per_step_at_mean_basis": {"j": jmean, **{k: round(v, 1) for k, v in per_step.items()}}}


def measured_traffic(kernel, N, batch=1):
    """HBM bytes per launch of `kernel` at grid size N and `batch` vectors per
    launch as measured by rocprofv3 (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00,
    copy-calibrated) and committed in profiles/hbm_traffic_n512.json (key
    `kernel`, or `kernel`_b<batch> for batches): the entry (bytes_per_launch,
    algorithmic, ...), or None when none matches."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_n512.json")
    key = kernel if batch == 1 else f"{kernel}_b{batch}"
    try:
        with open(path) as f:
            t = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    if not t or t.get("N") != N:
        return None
    return t


def profile_fno(ctx, net, dchi, db, batch=1):
    """Per-kernel durations of one forward pass of the network on `batch`
    vectors (HIP events on the context stream, cgm_fno_profile) with their
    algorithmic bytes."""
    import torch
    from cugmres_amd import _lib
    from cugmres_amd._lib import lib
    N = ctx.N
    n = N * N
    dchi = dchi.reshape(batch, n).contiguous()
    db = db.reshape(batch, n)
    v = (db / torch.linalg.vector_norm(db, dim=1, keepdim=True)).contiguous()
    z = torch.empty_like(v)
    ms = (ctypes.c_float * 9)()
    st = lib().cgm_fno_profile(net.handle, ctypes.c_void_p(dchi.data_ptr()), ctypes.c_void_p(v.data_ptr()),
                               ctypes.c_void_p(z.data_ptr()), batch, 50, ms)
    _lib.check(st, "cgm_fno_profile")
    names = ["fno_lift", "fno_rowproj", "fno_dftx", "fno_mix", "fno_invx", "fno_cell_gemm", "fno_proj",
             "fno_forward_total", "fno_cell_gemm_lift"]
    # launches per forward pass: fno_lift = k_lift_row / k_lift_rows (lift +
    # layer-0 row projection; h0 stays in LDS); fno_cell_gemm = k_cell_row /
    # k_cell_rows (B_x of the row, cell GEMM + GELU + next row projection,
    # layers 1-2); fno_cell_gemm_lift = the same for layer 0 (k_cell_row /
    # k_cell_rows_lift), lifting its row from chi and v; fno_proj =
    # k_cell_proj (layer-3 cell GEMM + the 32 -> 2 projection).  Batches of
    # at least 2048 grid rows take the *_rows kernels, several grid rows per
    # work-group; the algorithmic bytes are the same.
    # For one vector the row kernels of
    # layers 0-2 form B_x themselves (inverse x-DFT from the mixed modes) and
    # k_invx runs for layer 3 only; for more, k_invx forms every B_x.  The
    # stand-alone row projection runs only in the backward pass.  Algorithmic
    # bytes per launch: BF16 activations 64 B per cell, complex64 vectors 8 B,
    # chi 4 B; FP32 row projections 4 KiB per grid row, x-DFT partials
    # 128 KiB per 64 rows, spectral weights 4 MiB per layer, mixed modes
    # 128 KiB per vector, row twiddles 256 B per row, BF16 per-row mixing
    # matrices 2 KiB per row
    # (per vector; the spectral weights once per launch)
    part = (N // 64) * 16384 * 8   # x-DFT partials: one per 64 rows
    Bv = batch
    bx = 128 * 1024 if batch == 1 else N * 2048   # B_x source per vector: modes (folded) or k_invx's rows
    bytes_ = {"fno_lift": Bv * (n * (4 + 8) + N * 4096), "fno_rowproj": Bv * (n * 64 + N * 4096),
              "fno_dftx": Bv * (N * 4096 + part) + N * 256, "fno_mix": Bv * (part + 16384 * 8) + 4 * 1024 * 1024,
              "fno_invx": Bv * (16384 * 8 + N * 2048) + N * 256,
              "fno_cell_gemm": Bv * (n * 128 + N * 4096 + bx) + N * 256,
              "fno_cell_gemm_lift": Bv * (n * (4 + 8 + 64) + N * 4096 + bx) + N * 256,
              "fno_proj": Bv * (n * (64 + 16) + N * 2048)}
    per_step = {"fno_lift": 1, "fno_rowproj": 0, "fno_dftx": 4, "fno_mix": 4, "fno_invx": 1 if batch == 1 else 4,
                "fno_cell_gemm": 2, "fno_cell_gemm_lift": 1, "fno_proj": 1}
    kernels = {}
    for i, k in enumerate(names):
        us = ms[i] * 1e3
        e = {"us": us}
        if k in bytes_:
            e.update(bytes=bytes_[k], GBps=bytes_[k] / (us * 1e-6) / 1e9, per_step=per_step[k])
        kernels[k] = e
    return {"kernels": kernels}


def profile_kernels(ctx, g, dchi, db, restart):
    """Per-pass durations of the FFT matvec and one Arnoldi step, timed with
    HIP events on the context stream (cgm_profile_step)."""
    import numpy as np
    import torch

    from cugmres_amd import _lib
    from cugmres_amd._lib import lib

    n = ctx.N * ctx.N
    names = ["pass1_row_fft", "pass2_col_fft_mul", "pass3_row_ifft", "cgs2_dots",
             "cgs2_update_dots", "cgs2_update_norm_givens"]
    ms = (ctypes.c_float * len(names))()
    reps = 50
    st = lib().cgm_profile_step(ctx.handle, ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(dchi.data_ptr()),
                                ctypes.c_void_p(db.data_ptr()), restart, reps, ms)
    _lib.check(st, "cgm_profile_step")
    us = {k: ms[i] * 1e3 for i, k in enumerate(names)}
    N, M = ctx.N, ctx.M
    # algorithmic bytes per launch (complex64 = 8 B, chi float = 4 B), quarter
    # spectrum (M/2+1)^2 * 8 B
    H = M // 2 + 1
    bytes_ = {
        "pass1_row_fft": 8 * N * N + 4 * N * N + 8 * N * M,
        "pass2_col_fft_mul": 8 * N * M * 2 + 8 * H * H,
        "pass3_row_ifft": 8 * N * M + 8 * N * N * 2,
    }
    nv = restart // 2 + 1  # mean basis size of a cycle
    bytes_["cgs2_dots"] = 8 * n * (nv + 1)
    bytes_["cgs2_update_dots"] = 8 * n * (nv + 2)
    bytes_["cgs2_update_norm_givens"] = 8 * n * (nv + 2)
    kernels = {k: {"us": us[k], "bytes": bytes_[k], "GBps": bytes_[k] / (us[k] * 1e-6) / 1e9}
               for k in names}
    t_mv = sum(us[k] for k in names[:3])
    b_mv = sum(bytes_[k] for k in names[:3])
    kernels["matvec_total"] = {"us": t_mv, "bytes": b_mv, "GBps": b_mv / (t_mv * 1e-6) / 1e9}
    dom = max(names, key=lambda k: us[k])
    a = kernels[dom]["GBps"]
    roof = {"bound": "hbm", "kernel": dom, "achieved": a, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": a / HBM_PEAK_GBS, "traffic": None,
            "per_unit": f"{kernels[dom]['bytes']} B per launch at N={N}, M={M}, basis {nv}"}
    return {"roofline": roof, "kernels": kernels}


def reference_problem(ctx, restart, reps=5):
    """The reference's benchmark problem at N = 512 (eps = 2.25 disk of
    radius N/6, k = 2*3.14/(N/6) per cell, x0 = b = plane wave along the row
    index, B/CUDA_GMRES/GMRES_kernels.cuh:12-19,557-572), solved to a true
    relative residual of 1e-6 with the same GMRES(restart).  The reference
    published 44.38 ms on a V100 for its m = 37 solve (Results/V100_mean_times.npy)."""
    import numpy as np
    import torch
    from cugmres_amd import solver
    N = ctx.N
    k = 2 * 3.14 / (N / 6)
    g = ctx.spectrum_reference(k, 0.5, 0.0)
    i, j = np.mgrid[0:N, 0:N]
    mask = ((i - N / 3) ** 2 + (j - N / 2) ** 2 <= (N / 6) ** 2)
    chi = ctx.to_device((solver.reference_chi(N) * mask).reshape(-1), "f32")
    b = ctx.to_device(np.repeat(np.exp(-1j * k * np.arange(N))[:, None], N, axis=1).reshape(-1))
    x = b.clone()
    ctx.gmres(g, chi, b, x, restart=restart, tol=1e-6, max_cycles=20)
    torch.cuda.synchronize()
    t = time.perf_counter()
    for _ in range(reps):
        x.copy_(b)
        _, info = ctx.gmres(g, chi, b, x, restart=restart, tol=1e-6, max_cycles=20)
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t) / reps
    return {"problem": "reference eps=2.25 disk, N=512, x0=b, tol 1e-6", "restart": restart,
            "iterations": info["iterations"], "converged": info["converged"],
            "rel_residual": info["rel_residual"], "ms_per_solve": dt * 1e3, "solves_per_s": 1.0 / dt,
            "V100_published_ms": 44.38, "speedup_vs_V100": 44.38e-3 / dt}


_CPU_WORKER = """
import json, sys, time
sys.path.insert(0, {root!r})
from cugmres_amd import problems
from oracle import oracle as orc
N, restart, seed = {N}, {restart}, {seed}
chi, b, _ = problems.config_plain(N, seed=seed)
op = orc.Op(N, orc.baseline_green_block(N, problems.KH))
t = orc.now()
op.gmres(chi.reshape(-1), b.reshape(-1), restart=restart, tol=1e-30, max_cycles=1, precision="c64")
print(json.dumps({{"cycle_s": orc.now() - t}}))
"""


def _cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(N, restart, budget_s, max_cycles):
    """The oracle's complex64 GMRES(restart) - the reference's MKL-path
    algorithm (B/MKL_GMRES_test_09122019/source/MKL_GMRES.h) restated in C -
    on the host cores, on the plain workload of the same seeds.  Like the GPU's
    plain GMRES it stalls on this physics, so its rate is quoted to the same
    cap (max_cycles restart cycles per system) and it reaches 1e-6 on none.

    One core: one restart cycle of seed 0 in this process.  All cores: one
    child process per core (up to 16), each timing one cycle of its own seed;
    the throughput is cores / slowest cycle.  The oracle's FFT is its own
    (mixed-radix, twiddles computed once per operator; no MKL on this image),
    where the reference rebuilds its DFTI descriptor on every matvec
    (B/MKL_GMRES_test_09122019/source/mkl_based_functions.h:5-15).  Also C0:
    the oracle's complex128 full GMRES (m <= 200) at N = 64 on seeds 0-9."""
    import subprocess
    from cugmres_amd import problems
    from oracle import oracle as orc
    chi, b, _ = problems.config_plain(N, seed=0)
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH))
    t = orc.now()
    op.matvec_c64(chi.reshape(-1), b.reshape(-1))
    t_mv = orc.now() - t
    t = orc.now()
    op.gmres(chi.reshape(-1), b.reshape(-1), restart=restart, tol=1e-30, max_cycles=1, precision="c64")
    t_cycle = orc.now() - t
    del op
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    procs = [subprocess.Popen([sys.executable, "-c", _CPU_WORKER.format(root=ROOT, N=N, restart=restart, seed=s)],
                              stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True) for s in range(cores)]
    t0 = time.perf_counter()
    cyc = []
    for p_ in procs:
        o, _ = p_.communicate(timeout=max(120.0, 20 * budget_s))
        try:
            cyc.append(json.loads(o.strip().splitlines()[-1])["cycle_s"])
        except (ValueError, IndexError, KeyError):
            pass
    wall_all = time.perf_counter() - t0
    all_rate = (len(cyc) / max(cyc)) if cyc else None   # cycles/s, the slowest worker's cycle
    # C0: full GMRES (no restart), m <= 200, complex128, N = 64, seeds 0..9
    c0 = []
    for s_ in range(10):
        chi0, b0, _ = problems.config_cpu(64, seed=s_)
        op0 = orc.Op(64, orc.baseline_green_block(64, problems.KH))
        t = orc.now()
        _, info = op0.gmres(chi0.reshape(-1), b0.reshape(-1), restart=200, tol=1e-6, max_cycles=1)
        c0.append((orc.now() - t, info["rel_residual"], info["converged"], info["iterations"]))
    return {"value": 1.0 / (t_cycle * max_cycles), "unit": f"capped solves/s (GMRES({restart}), {max_cycles} "
                                                            f"cycles; none reach 1e-6: 0.0 solves/s to 1e-6)",
            "cores": 1, "kind": "port",
            "sample": f"1 GMRES({restart}) cycle at N={N} c64, seed 0: {t_cycle:.2f} s (matvec {t_mv * 1e3:.0f} ms)",
            "all_cores": {"cores": cores, "value": (all_rate / max_cycles) if all_rate else None,
                          "cycles": len(cyc), "wall_s": round(wall_all, 2)},
            "cpu_model": _cpu_model(), "nproc": os.cpu_count(), "fft": "oracle's own, twiddles cached",
            "C0_N64_c128_m200": {"seeds": 10, "s_per_solve": _stats([c[0] for c in c0]),
                                 "rel_residual": _stats([c[1] for c in c0]),
                                 "converged": int(sum(c[2] for c in c0)), "iterations": _stats([c[3] for c in c0])}}


if __name__ == "__main__":
    main()
