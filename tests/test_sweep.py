"""The benchmark-sweep driver: the reference's problem, its statistics and its
output files ([RPW] Timestamps & Speedups.ipynb cell 0; B/CUDA_GMRES/launch_GMRES.cuh:6-239)."""
import hashlib
import json
import math
import os

import numpy as np
import pytest

from cugmres_amd import formats, sweep

GOLDEN = os.path.join(os.path.dirname(__file__), "golden")


@pytest.mark.parametrize("N", [32, 64, 128, 256, 512, 1024])
def test_mask_is_the_reference_cylinder(N):
    # bit-exact against the sha256 of the reference's cylinder_N.txt
    with open(os.path.join(GOLDEN, "masks.json")) as f:
        info = json.load(f)[str(N)]
    m = sweep.reference_mask(N)
    assert int(m.sum()) == info["ones"]
    assert hashlib.sha256(formats.mask_text(m).encode()).hexdigest() == info["sha256"]


def test_refined_mean_drops_outliers():
    # mean 22, population std sqrt(1522): only 100 lies a std or more away
    mean, sd = sweep.refined([1.0, 2.0, 3.0, 4.0, 100.0])
    assert sd == pytest.approx(math.sqrt(1522.0))
    assert mean == pytest.approx(2.5)


def test_refined_keeps_within_strict_band():
    # the band is open: both samples sit exactly one std from the mean, none is
    # kept and the plain mean is reported
    mean, sd = sweep.refined([0.0, 2.0])
    assert (mean, sd) == (1.0, 1.0)
    mean, sd = sweep.refined([5.0] * 4)
    assert (mean, sd) == (5.0, 0.0)


def test_published_table_is_consistent():
    # every device has the same 12 (N, maxiter) entries, times grow with N and
    # with the Krylov dimension, and the V100 is the fastest device at each
    keys = set(sweep.PUBLISHED_MS["V100"])
    assert len(keys) == 12
    for dev, t in sweep.PUBLISHED_MS.items():
        assert set(t) == keys
        for N in (256, 512, 1024, 2048):
            assert t[(N, 3)][0] < t[(N, 30)][0] < t[(N, 49)][0]
        for m in (3, 30, 49):
            assert all(t[(a, m)][0] < t[(b, m)][0] for a, b in [(256, 512), (512, 1024), (1024, 2048)])
    for k in keys:
        assert sweep.PUBLISHED_MS["V100"][k][0] <= sweep.PUBLISHED_MS["GTX1080Ti"][k][0]
        assert sweep.PUBLISHED_MS["GTX1080Ti"][k][0] < sweep.PUBLISHED_MS["MKL"][k][0]


@pytest.mark.gpu
def test_sweep_writes_reference_files(tmp_path):
    from cugmres_amd.solver import n_timestamps
    r = sweep.run(Ns=(64,), maxiters=(3, 5), reps=2, out=str(tmp_path), analytical=True)
    for m in (3, 5):
        v = r[(64, m)]
        assert v["reps"] == 2 and v["steps"] == m and v["ms_mean"] > 0
        assert 0 < v["final_residual"] < 1e3
        for rep in range(2):
            name = lambda what: formats.result_name(str(tmp_path), 64, what, m, rep)
            res = formats.read_floats(name("residual"))
            assert res.shape == (1,)
            secs = formats.read_floats(name("maxiter"))
            assert secs.shape == (1,) and 0 < secs[0] < 60
            err = formats.read_floats(name("relative_error"))
            assert err.shape == (1,) and 0 < err[0] < 2
            with open(name("times")) as f:
                ns = [int(line) for line in f if line.strip()]
            assert len(ns) == n_timestamps(m)
            assert all(b >= a for a, b in zip(ns, ns[1:]))
        assert res[0] == pytest.approx(v["final_residual"], rel=1e-6)
