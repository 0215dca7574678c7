"""The benchmark's defaults and the committed measurements they rely on (CPU).

bench.py reports the dominant kernel's HBM traffic from
profiles/hbm_traffic_n512.json; the default configuration must find its
entry there, or the reported `traffic` would silently be null."""
import importlib.util
import os
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.fixture(scope="module")
def bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", os.path.join(ROOT, "bench.py"))
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod


def _defaults(bench, monkeypatch, *args):
    monkeypatch.setattr(sys, "argv", ["bench.py", *args])
    return bench.parse()


def test_defaults_are_the_benchmark_configuration(bench, monkeypatch):
    a = _defaults(bench, monkeypatch)
    assert (a.gpus, a.N, a.restart, a.tol, a.max_cycles) == (1, 512, 30, 1e-6, 100)
    assert (a.batch, a.steps, a.warmup) == (16, 1, 1)
    assert a.train_samples == 32000 and a.train_N == 256
    assert a.seed_offset == 0 and a.mode == "bench"


@pytest.mark.parametrize("batch", [1, 4, 16])
def test_traffic_is_committed_for_the_dominant_kernel(bench, batch):
    t = bench.measured_traffic("fno_cell_gemm", 512, batch)
    assert t is not None
    # 128 B per cell + 6 KiB per row per vector, measured within a few percent
    alg = batch * (512 * 512 * 128 + 512 * 4096) + 512 * 256 + batch * (128 * 1024 if batch == 1 else 512 * 2048)
    assert abs(t - alg) <= 0.08 * alg


def test_traffic_of_unmeasured_configurations_is_null(bench):
    assert bench.measured_traffic("fno_cell_gemm", 256, 16) is None
    assert bench.measured_traffic("no_such_kernel", 512, 1) is None
