"""bench.py contract pieces that run without a GPU: the multi-GPU refusal
(one GPU per box; the frame would shard by path-id slices, not built) and the
committed counter summaries that the bench line's roofline reads."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_committed_counter_summaries_parse():
    for k in ("trace", "shadow", "shade"):
        t = bench.measured_traffic(k)
        assert t is not None and t > 0, k
        busy, lanes = bench.measured_valu(k)
        assert 0.0 < busy <= 1.0 and 0.0 < lanes <= 1.0, (k, busy, lanes)
    assert bench.measured_traffic("raygen") is None


def test_roofline_is_against_the_binding_resource():
    """With the committed counter summaries and the launch times those runs
    saw (about 5.08 / 2.13 / 1.29 ms per trace / shadow / shade launch), the
    line's roofline is the traversal's VALU issue with a fraction in (0, 1],
    and every per-kernel HBM fraction (calibrated counter bytes over launch
    time against 8 TB/s) is in (0, 1]."""
    st = {"trace": {"ms": 5.08 * 8, "launches": 8}, "shadow": {"ms": 2.13 * 8, "launches": 8},
          "shade": {"ms": 1.29 * 8, "launches": 8}}
    bytes_step = {"trace": 6.0e11, "shadow": 1.5e11, "shade": 8.0e10}
    r = bench.frame_roofline("trace", st, 1, bytes_step, 208_297_870)
    assert r["bound"] == "valu" and r["unit"] == "Tlane-op/s"
    assert 0.0 < r["frac"] <= 1.0 and abs(r["frac"] - r["achieved"] / r["peak"]) < 1e-3
    assert abs(r["peak"] - 157.3 / 2) < 0.1      # the published FP32 vector peak in lane-FMA/s
    assert r["traffic"] == bench.measured_traffic("trace")
    for k, v in r["hbm"].items():
        assert 0.0 < v["frac"] <= 1.0, k
    assert 1000 < r["valu_issue"]["lane_ops_per_segment"] < 10000


def test_host_cpu_description():
    h = bench.host_cpu()
    assert h["usable"] >= 1 and h["usable"] <= h["affinity"] <= (h["cpu_count"] or h["affinity"])
