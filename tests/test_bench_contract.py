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
