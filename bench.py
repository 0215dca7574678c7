"""Benchmark of the GMRES solve on one MI355X: learned-FGMRES vs plain GMRES.

Workload (one "step" = one batch of --batch systems solved in lockstep with
each method): the plain GMRES workload at N = 512 (262,144 unknowns,
1024 x 1024 FFT embedding), complex64, tolerance 1e-6 on the true relative
residual, x0 = 0.  The permittivity map is a seeded binary map of 8 x 8-cell
blocks with values {1, 12.1} over the central 75 % of the area, h = lambda/20,
and the plane wave comes in at a seeded angle; every system has its own seed.
Step k solves seeds k*B .. k*B+B-1 with
  - plain GMRES(30) (cgm_gmres_batched: fused lagged-CGS2 kernels, every
    launch covering the B systems), and
  - flexible GMRES(30) right-preconditioned by the FNO network
    (cgm_fgmres_batched: one network forward pass per step for the B
    systems), trained on the GPU at N = 256 before the timed region (or
    loaded with --fno-params).
Each system's iterates equal its own single solve (tested), and the same B
systems are also solved one at a time ("single").  All inputs are generated
and copied to HBM before the timed region.

The default batch of 16 is one step of 16 systems (seeds 16..31 after a
warm-up step on seeds 0..15).

A solve stops at the tolerance or after --max-cycles restart cycles; the
number of converged solves, the achieved residuals and the wall time each
method needs to reach a common residual level are printed, so a capped run
is never reported as converged.

Printed: one JSON line with the metric (value = learned-FGMRES solves/s), the
plain GMRES figures beside it, the roofline of the learned path's dominant
kernel (HIP events on the solver's stream) and the CPU baseline (the oracle's
complex64 GMRES(30) restatement timed on the host).
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
HBM_PEAK_GBS = 8000.0


def parse():
    p = argparse.ArgumentParser()
    p.add_argument("--gpus", type=int, default=1)
    p.add_argument("--steps", type=int, default=1)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--N", type=int, default=512)
    p.add_argument("--restart", type=int, default=30)
    p.add_argument("--tol", type=float, default=1e-6)
    p.add_argument("--max-cycles", type=int, default=100)
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="budget of the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--fno-params", default=None, help="trained network (.npz from cugmres_amd.learn)")
    p.add_argument("--train-samples", type=int, default=32000,
                   help="training samples at --train-N; 32,000 was chosen on held-out systems "
                        "(--seed-offset 1000): longer training fits N = 256 at the expense of N = 512")
    p.add_argument("--train-N", type=int, default=256)
    p.add_argument("--seed-offset", type=int, default=0,
                   help="first seed of the solved systems (validation runs use seeds disjoint from the "
                        "default 0 .. (warmup + steps) * batch - 1)")
    p.add_argument("--stall-ratio", type=float, default=0.0,
                   help="FGMRES falls back to plain GMRES when a cycle leaves more than this "
                        "fraction of the residual (0: never)")
    p.add_argument("--plain-only", action="store_true")
    p.add_argument("--batch", type=int, default=16,
                   help="systems per lockstep batch (one step); 16 as the 16 right-hand sides per map of "
                        "batched workloads: the network's small spectral kernels are amortised over more "
                        "vectors (learned 3.4 / 3.8 / 4.0 capped solves/s at 4 / 8 / 16)")
    p.add_argument("--mode", choices=["bench", "step", "copy", "batched", "fno-step", "copybw", "sweep", "cycles"],
                   default="bench",
                   help="bench: the benchmark; step: only the per-kernel profile of one Arnoldi "
                        "step (for rocprofv3 --pmc); copy: 3 copies of 256 MiB (counter calibration); "
                        "batched: lockstep batches of 16 right-hand sides vs single solves; fno-step: "
                        "the per-kernel profile of one network forward pass on --batch vectors (for rocprofv3 "
                        "--pmc); copybw: "
                        "device-copy bandwidth at 8-256 MB; sweep: the reference's own benchmark "
                        "(its cylinder problem at N = 256..2048, Krylov dimension 3/30/49) through "
                        "cgm_Fast_GMRES_with_CUDA beside its published V100 / GTX 1080 Ti times; cycles: "
                        "3 x (2 restart cycles of cgm_fgmres_batched with a seeded untrained network, then "
                        "2 of cgm_gmres_batched) on --batch systems (for rocprofv3 kernel traces and "
                        "counters of whole solver steps)")
    p.add_argument("--sweep-reps", type=int, default=20)
    return p.parse_args()


def main():
    a = parse()
    if a.gpus != 1:
        print(json.dumps({"error": "this build measures one GPU; --gpus %d is not supported "
                                   "(the workload does not shard across GPUs)" % a.gpus}))
        sys.exit(2)
    import numpy as np
    import torch

    from cugmres_amd import problems, solver
    from cugmres_amd._lib import lib

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
    g = ctx.spectrum_helmholtz(problems.KH)
    B = a.batch
    if B < 1 or (B > 1 and a.stall_ratio > 0):
        print(json.dumps({"error": "--batch must be >= 1, and batched learned solves have no stall fallback "
                                   "(use --stall-ratio 0 or --batch 1)"}))
        sys.exit(2)
    inputs = []  # per system: (chi, b, theta) on the device
    for s_ in range((a.warmup + a.steps) * B):
        chi, b, theta = problems.config_plain(N, seed=a.seed_offset + s_)
        inputs.append((ctx.to_device(chi.reshape(-1), "f32"), ctx.to_device(b.reshape(-1)), theta))
    batches = [(torch.stack([inputs[k * B + j][0] for j in range(B)]),
                torch.stack([inputs[k * B + j][1] for j in range(B)])) for k in range(a.warmup + a.steps)]
    xb = torch.zeros(B, N * N, dtype=torch.complex64, device=ctx.device)
    x = torch.zeros(N * N, dtype=torch.complex64, device=ctx.device)

    # ---- the learned preconditioner (trained before the timed region) -------
    net, training = None, None
    if not a.plain_only:
        from cugmres_amd import fno, learn
        if a.fno_params:
            params = fno.load(a.fno_params)
            training = {"loaded": a.fno_params}
        else:
            tctx = solver.Context(a.train_N, 0)
            t = time.perf_counter()
            tnet, hist = learn.train(tctx, samples=a.train_samples, batch=8, lr=1e-3, seed=0)
            torch.cuda.synchronize()
            params = tnet.params.cpu().numpy()
            training = {"N": a.train_N, "samples": a.train_samples, "batch": 8, "seconds": time.perf_counter() - t,
                        "loss_first": hist[0][1], "loss_last_100_mean": float(np.mean([h[1] for h in hist[-100:]])),
                        "kernels": "cgm_fno_forward/backward (BF16 MFMA)"}
            tnet.close()
            tctx.close()
        net = fno.Preconditioner(ctx, batch=B, params=params)
    torch.cuda.synchronize()

    def plain(k):
        chib, bb = batches[k]
        return ctx.gmres_batched(g, chib, bb, xb.zero_(), restart=a.restart, tol=a.tol, max_cycles=a.max_cycles,
                                 history=True)[1]

    def learned(k):
        chib, bb = batches[k]
        if B == 1:
            return [ctx.fgmres(g, chib[0], bb[0], xb[0].zero_(), precond=net, restart=a.restart, tol=a.tol,
                               max_cycles=a.max_cycles, stall_ratio=a.stall_ratio, history=True)[1]]
        return ctx.fgmres_batched(g, chib, bb, xb.zero_(), precond=net, restart=a.restart, tol=a.tol,
                                  max_cycles=a.max_cycles, history=True)[1]

    def run(fn):
        for k in range(a.warmup):
            fn(k)
        torch.cuda.synchronize()
        infos = []
        t0 = time.perf_counter()
        for k in range(a.warmup, a.warmup + a.steps):
            infos.extend(fn(k))
        torch.cuda.synchronize()
        return infos, time.perf_counter() - t0

    p_infos, p_wall = run(plain)
    l_infos, l_wall = (run(learned) if net is not None else (None, None))

    def summary(infos, wall, nsolve):
        its = [i["iterations"] for i in infos]
        return {"solves_per_s": nsolve / wall, "ms_per_solve": wall / nsolve * 1e3,
                "converged": int(sum(i["converged"] for i in infos)), "of": len(infos),
                "iterations": its, "achieved_rel_residual": [i["rel_residual"] for i in infos],
                "us_per_arnoldi_step": wall / max(1, sum(its)) * 1e6}

    plain_s = summary(p_infos, p_wall, a.steps * B)
    plain_s["us_per_batch_step"] = plain_s["us_per_arnoldi_step"] * B
    out_learned = summary(l_infos, l_wall, a.steps * B) if l_infos else None
    if out_learned:
        out_learned["us_per_batch_step"] = out_learned["us_per_arnoldi_step"] * B
        out_learned["precond_cycles"] = [i["precond_cycles"] for i in l_infos]
        out_learned["fell_back"] = [i["fell_back"] for i in l_infos]
        # wall time to a common residual: the worst residual plain GMRES ends at
        level = max(plain_s["achieved_rel_residual"])
        def t_to(infos, us_per_it):
            res = []
            for i in infos:
                h = i["history"] / i["bnorm"]
                hit = np.nonzero(h <= level)[0]
                res.append(float(hit[0]) * us_per_it * 1e-3 if hit.size else None)
            return res
        tp = t_to(p_infos, plain_s["us_per_arnoldi_step"])
        tl = t_to(l_infos, out_learned["us_per_arnoldi_step"])
        ok = [(u, v) for u, v in zip(tp, tl) if u is not None and v is not None]
        matched = {"level": level, "plain_ms": tp, "learned_ms": tl,
                   "speedup": (sum(u for u, _ in ok) / sum(v for _, v in ok)) if ok else None}
    else:
        matched = None

    # ---- the same systems one at a time (first timed batch) -----------------
    def single(fn):
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        infos = [fn(j) for j in range(B)]
        torch.cuda.synchronize()
        w = time.perf_counter() - t0
        its = sum(i["iterations"] for i in infos)
        return {"solves_per_s": B / w, "ms_per_solve": w / B * 1e3, "us_per_arnoldi_step": w / max(1, its) * 1e6}

    def plain1(j):
        dchi, db, _ = inputs[a.warmup * B + j]
        return ctx.gmres(g, dchi, db, x.zero_(), restart=a.restart, tol=a.tol, max_cycles=a.max_cycles)[1]

    def learned1(j):
        dchi, db, _ = inputs[a.warmup * B + j]
        return ctx.fgmres(g, dchi, db, x.zero_(), precond=net, restart=a.restart, tol=a.tol,
                          max_cycles=a.max_cycles, stall_ratio=a.stall_ratio)[1]

    single_out = {"systems": B, "plain": single(plain1), "learned": single(learned1) if net is not None else None}

    # ---- roofline of the dominant kernel of the timed path -------------------
    prof = profile_kernels(ctx, g, inputs[0][0], inputs[0][1], a.restart)
    fprof = profile_fno(ctx, net, batches[0][0], batches[0][1], B) if net is not None else None
    roof = prof["roofline"]
    roof["traffic"] = measured_traffic(roof["kernel"], N)
    if fprof is not None:
        # per batch step of learned FGMRES: one forward pass for the B vectors (lift, 4 x x-DFT /
        # mixing / inverse, 3 row cell GEMMs, the projecting cell GEMM), then passes 1-2 of the B
        # matvecs of z_j and the fused flexible k_arn_a / k_arn_b; the heaviest kernel class by
        # time (the matvec and the fused pair are known only as the remainder of the step time)
        per_step = {k: v["us"] * v["per_step"] for k, v in fprof["kernels"].items() if "per_step" in v}
        dom = max(per_step, key=per_step.get)
        per_step["matvec_arnoldi_and_gaps"] = out_learned["us_per_batch_step"] - sum(per_step.values())
        src = fprof["kernels"]
        ach = src[dom]["GBps"]
        traffic = measured_traffic(dom, N, B)
        roof = {"bound": "hbm", "kernel": dom, "achieved": ach, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                "frac": ach / HBM_PEAK_GBS, "traffic": traffic,
                "per_unit": f"{src[dom]['bytes']} B per launch ({B} vectors at N={N})",
                "us_per_launch": src[dom]["us"], "us_per_batch_step": per_step}

    # ---- the reference's own problem at the same N ------------------------
    refp = reference_problem(ctx, a.restart) if N == 512 else None

    # ---- CPU baseline ------------------------------------------------------
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(N, a.restart, a.cpu_seconds, np.mean([i["cycles"] for i in p_infos]))

    head = out_learned if out_learned else plain_s
    out = {
        "metric": METRIC,
        "value": head["solves_per_s"],
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": head["ms_per_solve"] * B,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded eps maps and plane waves)",
        "config": {"workload": f"{'learned FGMRES' if out_learned else 'plain GMRES'}({a.restart}) N={N}, "
                               f"8x8 blocks eps in {{1,12.1}} over 75% area, h=lambda/20, tol {a.tol:g}, "
                               f"cap {a.max_cycles} cycles, {B} seeded systems per step in lockstep",
                   "N": N, "unknowns": N * N, "embedding": ctx.M, "restart": a.restart,
                   "tol": a.tol, "max_cycles": a.max_cycles, "batch": B},
        "learned": out_learned,
        "plain": plain_s,
        "speedup_solves_per_s": (out_learned["solves_per_s"] / plain_s["solves_per_s"]) if out_learned else None,
        "residual_matched": matched,
        "single": single_out,
        "training": training,
        "roofline": roof,
        "kernels": prof["kernels"],
        "fno_kernels": fprof["kernels"] if fprof else None,
        "cpu_baseline": cpu,
        "reference_problem": refp,
    }
    print(json.dumps(out), flush=True)


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


def measured_traffic(kernel, N, batch=1):
    """HBM bytes per launch of `kernel` at grid size N and `batch` vectors per
    launch as measured by rocprofv3 (FETCH_SIZE x 2.00 + WRITE_SIZE x 1.00,
    copy-calibrated) and committed in profiles/hbm_traffic_n512.json (key
    `kernel`, or `kernel`_b<batch> for batches); None when none matches."""
    path = os.path.join(ROOT, "profiles", "hbm_traffic_n512.json")
    key = kernel if batch == 1 else f"{kernel}_b{batch}"
    try:
        with open(path) as f:
            t = json.load(f).get(key)
    except (OSError, ValueError):
        return None
    if not t or t.get("N") != N:
        return None
    return t["bytes_per_launch"]


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


def cpu_baseline(N, restart, budget_s, cycles_per_solve):
    """The oracle's complex64 GMRES(restart) (the reference's MKL-path
    algorithm restated in C, one thread) timed on one restart cycle of the
    same N = 512 input; projected to solves/s with the GPU's cycles/solve."""
    from cugmres_amd import problems
    from oracle import oracle as orc
    chi, b, _ = problems.config_plain(N, seed=0)
    op = orc.Op(N, orc.baseline_green_block(N, problems.KH))
    t = orc.now()
    op.matvec_c64(chi.reshape(-1), b.reshape(-1))
    t_mv = orc.now() - t
    # a restart cycle costs restart+1 matvecs plus orthogonalisation; run
    # whole cycles while the budget lasts (at least one)
    t = orc.now()
    _, info = op.gmres(chi.reshape(-1), b.reshape(-1), restart=restart, tol=1e-30, max_cycles=1,
                       precision="c64")
    t_cycle = orc.now() - t
    per_solve = t_cycle * max(1.0, cycles_per_solve)
    return {"value": 1.0 / per_solve, "unit": "solves/s", "cores": 1, "kind": "port",
            "sample": f"1 GMRES({restart}) restart cycle at N={N} c64 ({t_cycle:.2f} s, matvec "
                      f"{t_mv * 1e3:.0f} ms), scaled by {cycles_per_solve:.1f} cycles/solve"}


if __name__ == "__main__":
    main()
