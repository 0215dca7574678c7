"""C-ABI library: loads without a GPU, exports every declared symbol, host
helpers agree with the oracle (CPU only)."""
import ctypes
import os
import re

import pytest

from conftest import ROOT, golden

HDR = os.path.join(ROOT, "include", "cugmres.h")


def declared_functions():
    txt = open(HDR).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    names = set(re.findall(r"\b(cgm_[A-Za-z0-9_]+)\s*\(", txt))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    from cugmres_amd import _lib
    L = _lib.lib()
    names = declared_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing
    assert set(names) <= set(_lib.SIGNATURES), set(names) - set(_lib.SIGNATURES)


def test_host_bessel_matches_scipy_fixture():
    from cugmres_amd import solver
    for x, j0, j1, y0, y1 in golden("hankel_scipy.json")["rows"][::7]:
        got = solver.bessel01(x)
        for g, r in zip(got, (j0, j1, y0, y1)):
            assert abs(g - r) <= 5e-13 * max(1.0, abs(r))


def test_reference_chi_matches_oracle(orc):
    from cugmres_amd import solver
    for N in [32, 64, 256, 512, 1024]:
        assert solver.reference_chi(N) == pytest.approx(orc.ref_chi(N), rel=0, abs=0)


def test_timestamp_schedule():
    """B/CUDA_GMRES/launch_GMRES.cuh:245-307; the reference's own N=256, m=44
    MKL run wrote exactly this many stamps (times_256_4400.txt)."""
    from cugmres_amd import solver
    g = golden("mkl_solution.npz")
    assert solver.n_timestamps(int(g["maxiter"])) == int(g["n_timestamps"])

    def direct(maxiter):
        n = 18
        gi = 1
        while gi < maxiter:
            n += 3 + 2 * (gi + 1) + 7
            gi += 1
        return n + 4 + gi + 2
    for m in range(0, 60):
        assert solver.n_timestamps(m) == direct(m)


def test_status_strings():
    from cugmres_amd import _lib
    L = _lib.lib()
    assert L.cgm_status_string(0) == b"success"
    assert L.cgm_status_string(4) == b"not converged"
    maj, mi = ctypes.c_int(), ctypes.c_int()
    L.cgm_version(ctypes.byref(maj), ctypes.byref(mi))
    assert (maj.value, mi.value) == (0, 1)


def test_create_fails_cleanly_without_gpu():
    import torch
    if torch.cuda.is_available():
        pytest.skip("a GPU is present")
    from cugmres_amd import _lib
    L = _lib.lib()
    p = ctypes.c_void_p()
    st = L.cgm_create(ctypes.byref(p), 64, 0)
    assert st == _lib.ERROR_DEVICE and not p.value
    assert b"device" in L.cgm_last_error().lower() or len(L.cgm_last_error()) > 0
