"""Benchmark of the GMRES solve on one MI355X.

Workload (one "step" = one solve): the plain GMRES workload at N = 512
(262,144 unknowns, 1024 x 1024 FFT embedding), complex64, restarted
GMRES(30), tolerance 1e-6 on the true relative residual, x0 = 0.  The
permittivity map is a seeded binary map of 8 x 8-cell blocks with values
{1, 12.1} over the central 75 % of the area, h = lambda/20, and the plane
wave comes in at a seeded angle.  Step k solves seed k.  All inputs are
generated and copied to HBM before the timed region.

A solve stops at the tolerance or after --max-cycles restart cycles; both
the number of converged solves and the achieved residuals are printed, so
a capped run is never reported as converged.

Printed: one JSON line (rank 0) with the metric, roofline of the dominant
kernel (measured live with HIP events on the solver's stream) and the CPU
baseline (the oracle's complex64 GMRES(30) restatement timed on the host).
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
    p.add_argument("--steps", type=int, default=4)
    p.add_argument("--warmup", type=int, default=1)
    p.add_argument("--N", type=int, default=512)
    p.add_argument("--restart", type=int, default=30)
    p.add_argument("--tol", type=float, default=1e-6)
    p.add_argument("--max-cycles", type=int, default=100)
    p.add_argument("--cpu-seconds", type=float, default=15.0,
                   help="budget of the CPU-baseline sample")
    p.add_argument("--no-cpu", action="store_true")
    p.add_argument("--mode", choices=["bench", "step", "copy"], default="bench",
                   help="bench: the benchmark; step: only the per-kernel profile of one Arnoldi "
                        "step (for rocprofv3 --pmc); copy: 3 copies of 256 MiB (counter calibration)")
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
    N = a.N
    ctx = solver.Context(N, 0)
    if a.mode == "step":
        chi, b, _ = problems.config_plain(N, seed=0)
        prof = profile_kernels(ctx, ctx.spectrum_helmholtz(problems.KH),
                               ctx.to_device(chi.reshape(-1), "f32"), ctx.to_device(b.reshape(-1)),
                               a.restart)
        print(json.dumps(prof))
        return
    stream = torch.cuda.current_stream()
    g = ctx.spectrum_helmholtz(problems.KH)
    seeds = list(range(a.warmup + a.steps))
    inputs = []
    for s in seeds:
        chi, b, theta = problems.config_plain(N, seed=s)
        inputs.append((ctx.to_device(chi.reshape(-1), "f32"), ctx.to_device(b.reshape(-1)), theta))
    x = torch.zeros(N * N, dtype=torch.complex64, device=ctx.device)
    torch.cuda.synchronize()

    def solve(k):
        dchi, db, _ = inputs[k]
        x.zero_()
        return ctx.gmres(g, dchi, db, x, restart=a.restart, tol=a.tol, max_cycles=a.max_cycles)[1]

    for k in range(a.warmup):
        solve(k)
    torch.cuda.synchronize()
    infos = []
    t0 = time.perf_counter()
    for k in range(a.warmup, a.warmup + a.steps):
        infos.append(solve(k))
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    wall = t1 - t0
    ms_per_step = wall / a.steps * 1e3
    iters = [i["iterations"] for i in infos]
    conv = [i["converged"] for i in infos]
    resid = [i["rel_residual"] for i in infos]

    # ---- roofline of the dominant kernel ---------------------------------
    prof = profile_kernels(ctx, g, inputs[0][0], inputs[0][1], a.restart)

    # ---- the reference's own problem at the same N ------------------------
    refp = reference_problem(ctx, a.restart) if N == 512 else None

    # ---- CPU baseline ------------------------------------------------------
    cpu = None
    if not a.no_cpu:
        cpu = cpu_baseline(N, a.restart, a.cpu_seconds, np.mean([i["cycles"] for i in infos]))

    out = {
        "metric": METRIC,
        "value": a.steps / wall,
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": a.steps,
        "warmup": a.warmup,
        "ms_per_step": ms_per_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "c64",
        "data": "synthetic (seeded eps maps and plane waves)",
        "config": {"workload": f"plain GMRES({a.restart}) N={N}, 8x8 blocks eps in {{1,12.1}} over "
                               f"75% area, h=lambda/20, tol {a.tol:g}, cap {a.max_cycles} cycles",
                   "N": N, "unknowns": N * N, "embedding": ctx.M, "restart": a.restart,
                   "tol": a.tol, "max_cycles": a.max_cycles},
        "solves": {"converged": int(sum(conv)), "of": len(conv), "iterations": iters,
                   "achieved_rel_residual": resid,
                   "us_per_arnoldi_step": wall / max(1, sum(iters)) * 1e6},
        "roofline": prof["roofline"],
        "kernels": prof["kernels"],
        "cpu_baseline": cpu,
        "reference_problem": refp,
    }
    print(json.dumps(out), flush=True)


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
