"""Backend benchmark harness (modisp_amd.benchmark): the reference's timing
protocol (tools/benchmark_backends.py: first call, 3 warm-up and 12 steady
calls per path in rotating order, median / quartiles, valid roots per second)
on the R64 / R32 paths, with device sampling and an environment record."""
import json
from types import SimpleNamespace

import numpy as np
import pytest

from modisp_amd import benchmark


def fake(valid, changes=()):
    calls = {"n": 0}

    def call():
        calls["n"] += 1
        status = np.ones((2, 3, 4), np.uint8)
        status.flat[:valid] = 0
        v = np.where(status == 0, 3., np.nan)
        if calls["n"] in changes:
            v = v + 1.
        return SimpleNamespace(status=status, phase_velocity=v, mode_mask=np.isfinite(v),
                               trace_flags=np.zeros((2, 4), np.uint32), search_level=np.zeros((2, 4), np.int32),
                               root_count=np.zeros((2, 4), np.int32), search_bounds=np.zeros((2, 2)))
    return call, calls


def test_protocol_counts_orders_and_summaries():
    f64, n64 = fake(10)
    f32, n32 = fake(10, changes=(7,))
    rec = benchmark.protocol({"R64": f64, "R32": f32})
    assert n64["n"] == n32["n"] == 1 + 3 + 12
    for p in ("R64", "R32"):
        r = rec[p]
        assert len(r["warmup"]) == 3 and len(r["steady"]) == 12
        s = r["summary"]
        assert s["steady_calls"] == 12 and s["valid_roots"] == 10 and s["q25_s"] <= s["median_s"] <= s["q75_s"]
        assert s["valid_roots_per_s"] == pytest.approx(10 / s["median_s"])
    assert rec["R64"]["summary"]["repeated_changes"] == 0 and rec["R32"]["summary"]["repeated_changes"] == 1
    first = [e["position"] for e in rec["R64"]["steady"]]
    assert first[:2] == [0, 1]                      # the path order rotates every round


def test_amd_smi_parsing_tolerates_both_layouts():
    one = {"gpu": 0, "usage": {"gfx_activity": {"value": 87, "unit": "%"}},
           "mem_usage": {"used_vram": {"value": 2048, "unit": "MB"}}, "power": {"socket_power": {"value": 612}},
           "temperature": {"hotspot": {"value": 71}}}
    for text in (json.dumps([one]), json.dumps({"gpu_data": [one]})):
        s = benchmark.parse_amd_smi(text)
        assert s["gpu_util"] == 87 and s["memory_used_bytes"] == 2048 * 2 ** 20 and s["power_w"] == 612
    assert benchmark.parse_amd_smi("not json") == {}
    assert benchmark.parse_amd_smi(json.dumps([{"gpu": 0}]))["gpu_util"] is None


def test_monitor_records_a_missing_tool():
    mon = benchmark.Monitor(interval=.01)
    mon.tool = None
    with mon:
        pass
    s = mon.summary()
    assert s["sample_count"] == 0 and s["tool"] is None


@pytest.mark.gpu
def test_r64_and_r32_rows_on_a_small_case(tmp_path):
    import modisp_amd
    if not modisp_amd.available(0):
        pytest.fail("no gfx950 device visible to the gpu tests")
    out = tmp_path / "bench.json"
    assert benchmark.main(["--case", "plain", "--models", "48", "--warmups", "1", "--rounds", "2",
                           "--output", str(out)]) == 0
    doc = json.loads(out.read_text())
    r64, r32 = doc["records"]["R64"]["summary"], doc["records"]["R32"]["summary"]
    assert r64["steady_calls"] == r32["steady_calls"] == 2
    assert r64["valid_roots"] == r32["valid_roots"] > 0 and r64["repeated_changes"] == r32["repeated_changes"] == 0
    assert doc["ratio_r32_over_r64"] > 0 and doc["environment"]["library_sha256"]
    assert doc["monitor"]["sample_count"] >= 1 or doc["monitor"]["errors"]
