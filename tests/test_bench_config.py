"""The benchmark's defaults and the shape of its output line (CPU).

The driver keeps about 8 KB of bench.py's output and parses its last JSON
line, so the line carries summaries only and must stay well under 4 KB at
any --steps; `value` counts only solves that reached the tolerance."""
import importlib.util
import json
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


def test_defaults_are_the_benchmark_configuration(bench):
    a = bench.parse([])
    assert (a.gpus, a.N, a.restart, a.tol, a.max_cycles) == (1, 512, 30, 1e-6, 100)
    assert (a.tile, a.lr_restart, a.train_N) == (16, 4096, 256)
    assert a.seed_offset == 0 and a.mode == "bench"


def _infos(k, converged):
    return [{"iterations": 1000 + i, "converged": converged(i), "rel_residual": 1e-6 * (1 + i % 3),
             "ms": 2500.0 + i, "seed": i, "history": list(range(50))} for i in range(k)]


def test_line_is_compact_at_320_systems(bench):
    """320 systems per method (the round-1 driver's --steps 20 with 16-system
    batches) still give one line far below 4 KB: no per-system arrays."""
    a = bench.parse(["--steps", "320"])
    learned = bench.summarize_method(_infos(320, lambda i: i % 5 != 0), 800.0, "learned")
    tiles = bench.summarize_method(_infos(320, lambda i: i % 7 != 0), 900.0, "tiles")
    plain = bench.summarize_method(_infos(320, lambda i: False), 50.0, "plain")
    roof = {"bound": "hbm", "achieved": 4000.0, "peak": 8000.0, "unit": "GB/s", "frac": 0.5, "traffic": None,
            "kernel": "k_lr_upd", "per_unit": "x" * 80, "us": 900.0, "us_per_step": {"lr_proj": 1.0}}
    cpu = {"value": 0.004, "unit": "capped solves/s", "cores": 1, "kind": "port", "sample": "y" * 120,
           "all_cores": {"cores": 16, "value": 0.05}, "cpu_model": "z" * 60, "nproc": 16,
           "C0_N64_c128_m200": {"seeds": 10, "s_per_solve": [1, 2, 3], "rel_residual": [1, 2, 3]}}
    out = bench.make_line(a, 512, 1024, learned, tiles, plain, 800.0, {"N": 256, "samples": 8000}, roof,
                          {"us": 20.0, "GBps": 1300.0}, cpu, {"problem": "reference", "ms_per_solve": 3.0})
    line = json.dumps(out, separators=(",", ":"))
    assert len(line) < 4096, len(line)
    back = json.loads(line)
    assert back["metric"] == bench.METRIC and back["unit"] == "solves/s"
    # only converged solves count
    assert back["value"] == pytest.approx(256 / 800.0)
    assert back["plain"]["solves_per_s"] == 0.0 and back["plain"]["converged"] == 0
    assert back["speedup_vs_plain"] is None
    assert back["learned"]["rel_residual"][0] == pytest.approx(1e-6)
    for k in ("roofline", "cpu_baseline"):
        assert isinstance(back[k], dict)


def test_traffic_lookup_of_unmeasured_kernels_is_null(bench):
    assert bench.measured_traffic("no_such_kernel", 512, 1) is None
