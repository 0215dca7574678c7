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
    assert (a.tile, a.lr_restart, a.train_N, a.train_samples) == (128, 8192, 512, 4000)
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
    live = {"proj": {"ms": 9000.0, "bytes": 5.2e13, "launches": 33000},
            "upd": {"ms": 9500.0, "bytes": 5.3e13, "launches": 32990}}
    roof = bench.live_roofline(live, 512, {"lr_proj": 1500.0, "lr_update": 1600.0, "tile_apply": 360.0,
                                            "matvec_x2": 41.0, "fno_forward": 103.0, "lr_reduce_givens": 30.0},
                               4300)
    cpu = {"value": 0.004, "unit": "capped solves/s", "cores": 1, "kind": "port", "sample": "y" * 120,
           "all_cores": {"cores": 16, "value": 0.05}, "cpu_model": "z" * 60, "nproc": 16,
           "C0_N64_c128_m200": {"seeds": 10, "s_per_solve": [1, 2, 3], "rel_residual": [1, 2, 3]}}
    out = bench.make_line(a, 512, 1024, learned, tiles, plain, 800.0, {"N": 256, "samples": 8000}, roof,
                          {"us": 19.65290028601885, "bytes": 26222600, "GBps": 1334.2865235343845,
                           "frac": 0.16678581544179807, "per_unit": "100 N^2 B per matvec (quarter spectrum)",
                           "batched16": {"us": 119.2, "frac": 0.391}},
                          cpu, {"problem": "reference", "ms_per_solve": 3.0})
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


def test_live_roofline_takes_the_longer_kernel_and_its_own_bytes(bench):
    """The dominant kernel is the one with the larger summed duration over the
    timed solves; achieved = its summed algorithmic bytes / summed time, and
    the mean basis size follows from its bytes per launch."""
    n = 512 * 512
    live = {"proj": {"ms": 1000.0, "bytes": 8.0 * n * (3000 + 2) * 400, "launches": 400},
            "upd": {"ms": 1200.0, "bytes": 8.0 * n * (3000 + 4) * 400, "launches": 400}}
    r = bench.live_roofline(live, 512, {"lr_proj": 1.0}, 3000)
    assert r["kernel"] == "k_lr_upd"
    # the tile application when it takes longer: 128 x 128 tiles, reduced form
    w = 16 * 6144 ** 2 * 4.0
    live["tile"] = {"ms": 1500.0, "bytes": 400 * (w + 16.0 * n), "launches": 400}
    rt = bench.live_roofline(live, 512, {"lr_proj": 1.0}, 3000, tile=128)
    assert rt["kernel"] == "k_tile_ws" and rt["frac"] == pytest.approx(rt["achieved"] / 8000.0)
    assert rt["achieved"] == pytest.approx(400 * (w + 16.0 * n) / 1.5 / 1e9)
    del live["tile"]
    assert r["achieved"] == pytest.approx(8.0 * n * 3004 * 400 / 1.2 / 1e9)
    assert r["frac"] == pytest.approx(r["achieved"] / 8000.0)
    assert r["us"] == pytest.approx(1200.0 * 1e3 / 400)
    assert "mean j 3000" in r["per_unit"]
    assert r["other"]["k_lr_proj"]["launches"] == 400

