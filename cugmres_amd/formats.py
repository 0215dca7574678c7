"""Readers/writers for the reference's on-disk formats.

  complex arrays   one value per line, "re im" (B/CUDA_GMRES/saveGPU.cuh:7-18,
                   get_array_C_to_CPU at saveGPU.cuh:181-199; Python reader
                   get_complex_array at B/PyGMRES/GMRES.py:71-92)
  masks            one 0/1 per line (B/CUDA_GMRES/launch_GMRES.cuh:57-66)
  float series     one value per line (residual, time and relative-error files,
                   B/CUDA_GMRES/launch_GMRES.cuh:184-218)
  timestamps       one integer nanosecond count per line
                   (B/MKL_GMRES_test_09122019/source/loop_functions.h:477-483)
Output files follow the reference's naming: <prefix>_<maxiter*100+rep>.txt.
"""
import os

import numpy as np


def read_complex(path, dtype=np.complex64):
    a = np.loadtxt(path, dtype=np.float64, ndmin=2)
    if a.size == 0:
        return np.zeros(0, dtype)
    return (a[:, 0] + 1j * a[:, 1]).astype(dtype)


def write_complex(path, a, digits=9):
    a = np.asarray(a).reshape(-1)
    fmt = f"%.{digits}g %.{digits}g"
    np.savetxt(path, np.column_stack([a.real, a.imag]), fmt=fmt)


def read_mask(path):
    with open(path) as f:
        vals = [1 if line.strip() == "1" else 0 for line in f if line.strip() != ""]
    return np.array(vals, np.uint8)


def write_mask(path, mask):
    m = np.asarray(mask).reshape(-1)
    with open(path, "w") as f:
        f.write("".join("1\n" if v else "0\n" for v in m))


def mask_text(mask):
    """The exact text of a mask file (used to compare with the reference's files)."""
    return "".join("1\n" if v else "0\n" for v in np.asarray(mask).reshape(-1))


def read_floats(path):
    return np.loadtxt(path, dtype=np.float64, ndmin=1)


def write_floats(path, a, digits=9):
    np.savetxt(path, np.asarray(a, np.float64).reshape(-1), fmt=f"%.{digits}g")


def write_timestamps(path, ns):
    with open(path, "w") as f:
        f.write("".join(f"{int(v)}\n" for v in np.asarray(ns).reshape(-1)))


def result_name(outdir, N, what, maxiter, rep):
    """output/time_N/<what>_<maxiter*100+rep>.txt (B/CUDA_GMRES/launch_GMRES.cuh:184-218)."""
    d = os.path.join(outdir, f"time_{N}")
    os.makedirs(d, exist_ok=True)
    return os.path.join(d, f"{what}_{maxiter * 100 + rep}.txt")
