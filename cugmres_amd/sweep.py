"""The reference's benchmark sweep on one GPU: its own problem, N x Krylov
dimension x repetitions, through the drop-in entry cgm_Fast_GMRES_with_CUDA.

Counterpart of the reference's driver (B/CUDA_GMRES/launch_GMRES.cuh:6-239:
cylinder mask, G_prep spectrum, x0 = b = exp(-i k i), a fixed number of
Arnoldi steps, the solution, residual and timestamp files per repetition) and
of the statistics of its notebook ([RPW] Timestamps & Speedups.ipynb, cell 0,
append_refined_statistics: the mean of the samples within one standard
deviation of the mean, beside the standard deviation of all of them).

The problem per N: eps = 2.25 disk of radius N/6 at (N/3, N/2) (the cylinder
mask of B/CUDA_GMRES/IN/cylinder_N.txt), k = 2*3.14/(N/6) per cell, chi =
(2.25 - 1) k^2 on the mask (B/CUDA_GMRES/GMRES_kernels.cuh:12-19), the
operator (i/4) H0(k sqrt((di - 1/2)^2 + dj^2)) of
B/MKL_GMRES_test_09122019/source/loop_functions.h:128-168 as a spectrum.
A solve is timed from the call to its return (synchronous), which includes
the callee-allocated residual array as the reference's timing does.

    python -m cugmres_amd.sweep --N 256 512 1024 2048 --maxiter 3 30 49 --reps 20 [--out DIR]
"""
import argparse
import ctypes
import json
import os
import time

import numpy as np

from . import formats
from ._lib import check, lib

# Published by the reference, (mean, std) in ms (Results/{V100,GTX1080Ti,
# MKL}_{mean,std}_times.npy, index [log2(N) - 8][maxiter - 3]; MKL is the
# sequential one-core CPU build)
PUBLISHED_MS = {
    "V100": {(256, 3): (2.13, 0.27), (256, 30): (20.40, 0.00), (256, 49): (43.96, 0.00),
             (512, 3): (3.53, 0.97), (512, 30): (33.18, 1.03), (512, 49): (66.58, 0.97),
             (1024, 3): (9.55, 4.34), (1024, 30): (100.04, 5.21), (1024, 49): (193.75, 4.79),
             (2048, 3): (42.54, 2.52), (2048, 30): (412.59, 5.00), (2048, 49): (769.40, 7.02)},
    "GTX1080Ti": {(256, 3): (2.18, 0.12), (256, 30): (26.19, 1.26), (256, 49): (55.10, 0.71),
                  (512, 3): (5.65, 0.52), (512, 30): (64.18, 1.40), (512, 49): (127.15, 0.71),
                  (1024, 3): (17.32, 0.41), (1024, 30): (190.76, 1.39), (1024, 49): (374.83, 2.80),
                  (2048, 3): (81.41, 17.46), (2048, 30): (808.91, 154.67), (2048, 49): (1516.01, 288.90)},
    "MKL": {(256, 3): (51.0, 23.0), (256, 30): (403.0, 34.0), (256, 49): (708.0, 64.0),
            (512, 3): (213.0, 19.0), (512, 30): (1761.0, 117.0), (512, 49): (3091.0, 219.0),
            (1024, 3): (955.0, 73.0), (1024, 30): (7880.0, 488.0), (1024, 49): (13677.0, 973.0),
            (2048, 3): (2745.0, 161.0), (2048, 30): (24254.0, 1134.0), (2048, 49): (44872.0, 2392.0)},
}


def reference_mask(N):
    """Cells with (i - N/3)^2 + (j - N/2)^2 <= (N/6)^2, the reference's cylinder."""
    i = np.arange(N, dtype=np.float64)[:, None]
    j = np.arange(N, dtype=np.float64)[None, :]
    return (i - N / 3.0) ** 2 + (j - N / 2.0) ** 2 <= (N / 6.0) ** 2


def refined(x):
    """(mean of the samples within one standard deviation of the mean, std of all)."""
    x = np.asarray(x, np.float64)
    sd = float(np.std(x))
    keep = np.abs(x - x.mean()) < sd
    return (float(x[keep].mean()) if keep.any() else float(x.mean())), sd


def run(Ns=(256, 512, 1024, 2048), maxiters=(3, 30, 49), reps=20, out=None, tol=1e-30, log=None,
        write_solution=False, analytical=False):
    """Returns {(N, maxiter): {...}}; with `out` also writes the reference's
    files out/time_N/{times,residual,maxiter}_{maxiter*100+rep}.txt (and
    solution_sample_* with write_solution).  analytical=True also reports
    (and writes as relative_error_*) ||x - u|| / ||u|| against the Mie field u
    of the same disk and wavenumber, on every max(1, N/256)-th row and column."""
    import torch
    from .solver import Context, n_timestamps
    results = {}
    for N in Ns:
        ctx = Context(N)
        k = 2.0 * 3.14 / (N / 6.0)
        g = ctx.spectrum_reference(k, 0.5, 0.0)
        mask = torch.as_tensor(reference_mask(N).reshape(-1)).to(ctx.device)
        x = torch.empty(N * N, dtype=torch.complex64, device=ctx.device)
        if analytical:
            from . import mie
            u_ref, rows, cols = mie.reference_problem_field(N, stride=max(1, N // 256), k=k)
        for m in maxiters:
            nts = n_timestamps(m)
            stamps = (ctypes.c_long * (2 * nts))()
            times, finals, errs = [], [], []
            for rep in range(-1, reps):              # one untimed warm-up
                check(lib().cgm_reference_plane_wave(ctx.handle, ctypes.c_void_p(x.data_ptr())),
                      "cgm_reference_plane_wave")
                ctx.synchronize()
                resid = ctypes.c_void_p()
                n_out = ctypes.c_uint()
                above = ctypes.c_bool()
                t0 = time.perf_counter()
                st = lib().cgm_Fast_GMRES_with_CUDA(
                    ctypes.c_void_p(g.data_ptr()), ctypes.c_void_p(mask.data_ptr()), ctypes.c_void_p(x.data_ptr()),
                    ctypes.byref(resid), ctypes.byref(n_out), ctx.handle, ctypes.c_float(tol), False, 0,
                    ctypes.byref(above), m, ctypes.cast(stamps, ctypes.c_void_p), N)
                t1 = time.perf_counter()
                check(st, "cgm_Fast_GMRES_with_CUDA")
                res = np.zeros(m + 1, np.float32)
                lib().cgm_memcpy_d2h(res.ctypes.data_as(ctypes.c_void_p), resid, (m + 1) * 4)
                lib().cgm_free(resid)
                if rep < 0:
                    continue
                times.append((t1 - t0) * 1e3)
                finals.append(float(res[n_out.value]))
                if analytical:
                    xs = x.cpu().numpy().reshape(N, N)[np.ix_(rows, cols)]
                    errs.append(float(np.linalg.norm(xs - u_ref) / np.linalg.norm(u_ref)))
                if out:
                    # the files of B/CUDA_GMRES/launch_GMRES.cuh:179-216: the
                    # final estimate, the elapsed seconds, the stamps (and the
                    # solution on request)
                    name = lambda what: formats.result_name(out, N, what, m, rep)
                    ns = np.frombuffer(stamps, np.int64).reshape(-1, 2)
                    formats.write_timestamps(name("times"), ns[:, 0] * 1_000_000_000 + ns[:, 1])
                    formats.write_floats(name("residual"), res[n_out.value:n_out.value + 1])
                    formats.write_floats(name("maxiter"), [t1 - t0])
                    if analytical:
                        formats.write_floats(name("relative_error"), errs[-1:])
                    if write_solution:
                        formats.write_complex(name("solution_sample"), x.cpu().numpy())
            mean, sd = refined(times)
            r = {"ms_mean": mean, "ms_std": sd, "reps": reps, "steps": m, "final_residual": finals[-1]}
            if analytical:
                r["relative_error"] = errs[-1]
            for dev, table in PUBLISHED_MS.items():
                if (N, m) in table:
                    r[f"{dev}_ms"] = table[(N, m)][0]
                    r[f"speedup_vs_{dev}"] = table[(N, m)][0] / mean
            results[(N, m)] = r
            if log:
                log(f"N={N} maxiter={m}: {mean:.3f} +- {sd:.3f} ms  " +
                    "  ".join(f"{d} {r.get(d + '_ms', float('nan')):.2f} ms" for d in PUBLISHED_MS))
        ctx.close()
    return results


def main():
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--N", type=int, nargs="+", default=[256, 512, 1024, 2048])
    ap.add_argument("--maxiter", type=int, nargs="+", default=[3, 30, 49])
    ap.add_argument("--reps", type=int, default=20)
    ap.add_argument("--out", default=None)
    ap.add_argument("--write-solution", action="store_true")
    ap.add_argument("--analytical", action="store_true", help="relative error against the Mie field")
    a = ap.parse_args()
    r = run(a.N, a.maxiter, a.reps, a.out, log=lambda s: print(s, flush=True), write_solution=a.write_solution,
            analytical=a.analytical)
    print(json.dumps({f"{N}x{m}": v for (N, m), v in r.items()}))


if __name__ == "__main__":
    main()
