"""Writes the golden fixtures of tests/golden/ from the reference's own data files.

Run in a checkout where the reference is mounted (read as data only: text via
numpy.loadtxt, .npy with allow_pickle=False).  The reference's source code is
never imported or executed.  Third-party values (scipy.special Bessel
functions) are written here too so the CPU tests do not depend on scipy.

    python tests/golden/make_fixtures.py [/root/reference]

Outputs (all small):
  gprep.npz            G_prep_32 and G_prep_64 in full, G_prep_128 strided
  masks.json           sha256 + ones-count of cylinder_{32..1024}.txt
  mkl_dumps.npz        N=256 MKL debug dumps: r0, A V0, V1, A V2 (strided
                       samples + full norms), H / Htemp / Jtotal (full)
  mkl_solution.npz     the N=256, m=44 MKL solution (full) and the
                       analytical (Mie) field (full), the logged residual,
                       the timestamp count
  python_histories.npz Results/residuals_* and relative_errors_* (N=256/512/1024)
  hankel_scipy.json    J0, J1, Y0, Y1 from scipy.special on a grid
"""
import hashlib
import json
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
B = os.path.join(ROOT, "Benchmarks of CUDA-based GMRES Solver for Toeplitz and Hankel Matrices "
                       "and Applications to Topology Optimization of Photonic Devices")
IN = os.path.join(B, "CUDA_GMRES", "Release", "input")
MKL = os.path.join(B, "MKL_GMRES_test_09122019")
RES = os.path.join(B, "Results")

STRIDE = 7  # prime stride for strided samples of N=256 vectors


def cplx(path):
    a = np.loadtxt(path, dtype=np.float64, ndmin=2)
    return (a[:, 0] + 1j * a[:, 1]).astype(np.complex64)


def sha(path):
    with open(path, "rb") as f:
        return hashlib.sha256(f.read()).hexdigest()


def main():
    out = {}
    g32 = cplx(os.path.join(IN, "G_prep_32.txt"))
    g64 = cplx(os.path.join(IN, "G_prep_64.txt"))
    g128 = cplx(os.path.join(IN, "G_prep_128.txt"))
    idx128 = np.arange(0, g128.size, 5)
    np.savez_compressed(os.path.join(HERE, "gprep.npz"), g32=g32, g64=g64,
                        g128_idx=idx128, g128_val=g128[idx128], g128_sha=sha(os.path.join(IN, "G_prep_128.txt")))

    masks = {}
    for N in [32, 64, 128, 256, 512, 1024]:
        p = os.path.join(IN, f"cylinder_{N}.txt")
        with open(p) as f:
            ones = sum(1 for line in f if line.strip() == "1")
        masks[str(N)] = {"sha256": sha(p), "ones": ones}
    with open(os.path.join(HERE, "masks.json"), "w") as f:
        json.dump(masks, f, indent=1)

    n = 256 * 256
    idx = np.arange(0, n, STRIDE)
    d = {"idx": idx}
    for key, fn in [("r0", "mkl_orthogonal_basis_2.txt"), ("r0b", "mkl_orthogonal_basis_3.txt"),
                    ("AV0", "mkl_orthogonal_basis_10.txt"), ("V1", "mkl_orthogonal_basis_1.txt"),
                    ("AV2", "mkl_orthogonal_basis_0.txt")]:
        v = cplx(os.path.join(MKL, fn))
        d[key] = v[idx]
        d[key + "_norm"] = np.linalg.norm(v.astype(np.complex128))
    d["H"] = cplx(os.path.join(MKL, "mkl_H__0.txt"))
    d["Htemp"] = cplx(os.path.join(MKL, "mkl_Htemp_0.txt"))
    d["Jtotal"] = cplx(os.path.join(MKL, "mkl_Jtotal_0.txt"))
    np.savez_compressed(os.path.join(HERE, "mkl_dumps.npz"), **d)

    sol = cplx(os.path.join(MKL, "_results", "mkl_solution_matrix.txt"))
    ana = cplx(os.path.join(IN, "analytical_solution_256.txt"))
    tr = os.path.join(MKL, "_results", "test_residuals")
    resid = float(np.loadtxt(os.path.join(tr, "residual_256_4400.txt")))
    tsec = float(np.loadtxt(os.path.join(tr, "time_256_4400.txt")))
    with open(os.path.join(tr, "times_256_4400.txt")) as f:
        nts = sum(1 for line in f if line.strip())
    ana64 = cplx(os.path.join(IN, "analytical_solution_64.txt"))
    np.savez_compressed(os.path.join(HERE, "mkl_solution.npz"), solution=sol, analytical=ana,
                        analytical_64=ana64, residual_estimate=resid, time_s=tsec,
                        n_timestamps=nts, maxiter=44)

    h = {}
    for N in [256, 512, 1024]:
        h[f"residuals_{N}"] = np.load(os.path.join(RES, f"residuals_{N}_0.npy"), allow_pickle=False)
        h[f"relative_errors_{N}"] = np.load(os.path.join(RES, f"relative_errors_{N}_0.npy"), allow_pickle=False)
    np.savez_compressed(os.path.join(HERE, "python_histories.npz"), **h)

    from scipy.special import jv, yv
    xs = np.concatenate([np.logspace(-3, np.log10(500.0), 400), [0.177245, 0.314159, 19.99, 20.0, 20.01]])
    rows = [[float(x), float(jv(0, x)), float(jv(1, x)), float(yv(0, x)), float(yv(1, x))] for x in xs]
    with open(os.path.join(HERE, "hankel_scipy.json"), "w") as f:
        json.dump({"source": "scipy.special.jv/yv", "rows": rows}, f)
    print("fixtures written to", HERE)


if __name__ == "__main__":
    main()
