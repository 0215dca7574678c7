"""The reference-named C++ shim (include/cugmres_compat.hpp) and a compiled
reference-style caller (tests/compat/caller.cpp, built by
__graft_entry__.build()).

The caller runs launch_GMRES's sequence (B/CUDA_GMRES/launch_GMRES.cuh:93-233)
through Fast_GMRES_with_CUDA, both G_x_fft_matvec overloads and cuSolve_LES
under their reference names (B/CUDA_GMRES/GMRES.cuh:22-35,450-458,
B/CUDA_GMRES/Fast_matvec.cuh:1-6,31-35); its outputs are checked against the
oracle.
"""
import os
import subprocess

import numpy as np
import pytest

from conftest import ROOT, golden, rel

CALLER = os.path.join(ROOT, "tests", "compat", "caller")


def test_compat_header_compiles_standalone(tmp_path):
    """The shim compiles as C++ on its own (no HIP headers needed) and
    declares the reference's names with their parameter lists."""
    src = tmp_path / "probe.cpp"
    src.write_text(
        '#include "cugmres_compat.hpp"\n'
        "void (*f1)(const cuComplex*, const bool*, cuComplex*, float**, unsigned int*, cufftHandle,\n"
        "           cublasHandle_t*, const float, const bool, const unsigned int, bool*, unsigned int,\n"
        "           cusolverDnHandle_t, timespec*, const unsigned int) = &Fast_GMRES_with_CUDA;\n"
        "void (*f2)(cuComplex*, bool*, cuComplex*, cuComplex*, cufftHandle, const unsigned int) =\n"
        "    &G_x_fft_matvec;\n"
        "void (*f3)(cuComplex*, cuComplex*, cuComplex*, cufftHandle, const unsigned int) = &G_x_fft_matvec;\n"
        "void (*f4)(cuComplex*, cuComplex*, const int, cusolverDnHandle_t, cuComplex*, int*, int*,\n"
        "           timespec*, unsigned int*) = &cuSolve_LES;\n"
        "static_assert(sizeof(cuComplex) == 8, \"layout of cuComplex\");\n")
    r = subprocess.run(["g++", "-std=c++17", "-Wall", "-Werror", "-fsyntax-only", "-I",
                        os.path.join(ROOT, "include"), str(src)], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_caller_is_built_and_links_the_library():
    assert os.path.exists(CALLER), "tests/compat/caller missing: run __graft_entry__.build()"
    out = subprocess.run(["ldd", CALLER], capture_output=True, text=True).stdout
    assert "libcugmres.so" in out and "not found" not in out.split("libcugmres.so")[1].split("\n")[0]


def _run_caller(orc, tmp_path, N, m):
    gprep = golden("gprep.npz")["g64"].astype(np.complex64)      # IN/G_prep_64.txt
    mask = orc.ref_cylinder_mask(N).reshape(-1).astype(np.uint8)
    gp = tmp_path / "gprep.bin"
    mk = tmp_path / "mask.bin"
    gprep.tofile(gp)
    mask.tofile(mk)
    prefix = str(tmp_path) + f"/m{m}_"
    print(f"compat caller N={N} maxiter={m}", flush=True)
    r = subprocess.run([CALLER, str(N), str(m), str(gp), str(mk), prefix], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    fields = dict(kv.split("=") for kv in r.stdout.split() if "=" in kv)
    blk, _ = orc.gprep_to_block(N, gprep)
    op = orc.Op(N, blk)
    chi = orc.ref_chi(N) * mask.astype(np.float64)
    b = orc.ref_plane_wave(N).reshape(-1).astype(np.complex128)
    return fields, (lambda name: np.fromfile(prefix + name, np.complex64)), op, chi, b


@pytest.mark.gpu
def test_reference_style_caller_matches_oracle(orc, tmp_path):
    """maxiter = 50, past the steep drop of this problem's residual between
    steps 20 and 29: the solution matches the oracle's complex128 GMRES to
    1e-5 (the oracle's own complex64 restatement sits 7e-7 from it); the
    for_gradient solution is within twice the oracle's own complex64-to-
    complex128 distance of the complex128 iterate."""
    N, m = 64, 50
    fields, load, op, chi, b = _run_caller(orc, tmp_path, N, m)
    assert int(fields["GMRES_n"]) == m and int(fields["grad_n"]) == m
    assert int(fields["clock_i"]) == 2
    assert int(fields["stamped"]) == int(fields["nts"])   # every timestamp of the schedule is written
    xr, info = op.gmres(chi, b, x0=b, restart=m, tol=0.0, max_cycles=1, history=True)
    idx = N * N // 2 + N // 3
    e = np.zeros(N * N, np.complex128)
    e[idx] = 1.0
    xg, ig = op.gmres(chi, e, x0=b, restart=m, tol=0.0, max_cycles=1, adjoint=True, history=True)
    # the gradient solve ends mid-convergence (residual 4e-5), where the
    # least-squares coefficients amplify rounding: the oracle's own complex64
    # restatement sits 3.1e-5 from its complex128 iterate
    xg32, _ = op.gmres(chi, e, x0=b, restart=m, tol=0.0, max_cycles=1, adjoint=True, history=True,
                       precision="c64")
    gap = rel(xg32, xg)
    A = np.zeros((4, 4), np.complex128)
    for rr in range(4):
        for c in range(4):
            A[rr, c] = (4.0 if rr == c else 0.0) + 0.25 * ((rr + 2 * c) % 3) + 0.125j * (rr - c)
    B = np.array([1.0 + rr - 0.5j * rr for rr in range(4)])
    got = {  # every comparison first, so one run reports all of them
        "residual0": (float(fields["residual0"]), info["beta0"]),
        "residual": (float(fields["residual"]), info["history"][m]),
        "solution": rel(load("solution.bin"), xr),
        "gradient": rel(load("gradient.bin"), xg),
        "grad_residual": (float(fields["grad_residual"]), ig["history"][m]),
        "matvec": rel(load("matvec.bin"), op.matvec(chi, b, 0)),
        "matvec_green": rel(load("matvec_green.bin"), op.matvec(chi, b, 2)),
        "les": rel(load("les.bin"), np.linalg.solve(A, B)),
    }
    print(got, flush=True)
    beta0 = got["residual0"][1]
    assert got["residual0"][0] == pytest.approx(beta0, rel=1e-5)
    assert got["matvec"] < 2e-6 and got["matvec_green"] < 2e-6
    assert got["les"] < 1e-6
    assert got["solution"] < 1e-5
    assert got["gradient"] < 1e-5 + 2 * gap
    # estimates at this depth are complex64 noise: within 1e-6 ||r0|| of the oracle
    assert abs(got["residual"][0] - got["residual"][1]) < 1e-6 * beta0
    assert got["grad_residual"][0] == pytest.approx(got["grad_residual"][1], rel=1e-3)
    assert int(fields["res_vs_tol"]) == int(got["residual"][0] > 0.001)


@pytest.mark.gpu
def test_reference_style_caller_tracks_complex64_oracle_at_the_drop(orc, tmp_path):
    """maxiter = 30 ends inside the drop, where single precision delays it: the
    oracle's complex64 restatement estimates 0.182 (complex128: 0.0027); the
    caller's estimate follows the complex64 one."""
    N, m = 64, 30
    fields, load, op, chi, b = _run_caller(orc, tmp_path, N, m)
    x64, i64 = op.gmres(chi, b, x0=b, restart=m, tol=0.0, max_cycles=1, history=True, precision="c64")
    print(float(fields["residual"]), i64["history"][m], flush=True)
    assert float(fields["residual"]) == pytest.approx(i64["history"][m], rel=0.1)
    assert rel(load("solution.bin"), x64) < 2e-3
