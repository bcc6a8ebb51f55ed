"""bench.py --vixen-json: the per-test results schema of VIXEN's benchmark runner
(libraries/Profiler/src/MetricsExporter.cpp:117-230, test ids :643-677) and the
fields tools/aggregate_results.py reads (configuration.pipeline, test_id, frames)."""
from __future__ import annotations

import importlib.util
import json
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parents[1]


def _bench():
    spec = importlib.util.spec_from_file_location("bench_under_test", ROOT / "bench.py")
    m = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(m)
    return m


def test_vixen_report_schema():
    b = _bench()
    ft = [0.30, 0.28, 0.29, 0.31]
    r = b.vixen_report(ft, 2073600, 5.3e8, 6.6e7, 14.2, 256, "tunnels", 1920, 1080, "AMD Instinct MI355X", "HIP 7.0",
                       288 * 1024.0, 1000.0, True, ["esvo"])
    json.dumps(r)  # serialisable
    assert r["test_id"] == "hip_256_TUNNELS_VX_RENDER_RUN1"
    assert set(r) == {"test_id", "timestamp", "configuration", "device", "frames", "statistics", "validation"}
    assert r["configuration"] == {"pipeline": "hip", "shader": "vx_render", "resolution": 256, "scene_type": "tunnels",
                                  "screen_width": 1920, "screen_height": 1080, "optimizations": ["esvo"]}
    assert set(r["device"]) == {"gpu", "driver", "vram_gb", "bandwidth_estimated"} and r["device"]["vram_gb"] == 288
    f0 = r["frames"][0]
    assert set(f0) == {"frame_num", "frame_time_ms", "fps", "bandwidth_read_gbps", "bandwidth_write_gbps",
                       "ray_throughput_mrays", "vram_mb", "avg_voxels_per_ray"}
    assert np.isclose(f0["ray_throughput_mrays"], 2073600 / 0.30e-3 / 1e6)
    assert np.isclose(f0["fps"], 1000 / 0.30) and np.isclose(f0["bandwidth_write_gbps"], 6.6e7 / 0.30e-3 / 1e9)
    st = r["statistics"]
    assert np.isclose(st["frame_time_mean"], np.mean(ft)) and np.isclose(st["frame_time_stddev"], np.std(ft))
    assert st["frame_time_p99"] <= max(ft) and st["frame_time_p99"] >= np.mean(ft)
    assert r["validation"] == {"valid": True, "error_count": 0, "warning_count": 0}
