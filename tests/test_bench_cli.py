"""bench.py's contract on the CPU: it measures one GPU, so --gpus N > 1 exits
(status 2) with a message before touching a device."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_bench_rejects_more_than_one_gpu():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2", "--steps", "1"],
                       capture_output=True, text=True, timeout=120, cwd=ROOT)
    assert r.returncode == 2, (r.returncode, r.stderr[-500:])
    assert "measures one GPU" in r.stderr and r.stdout == ""


def test_traversal_traffic_profile_has_what_bench_reads():
    """bench.py's roofline multiplies the profiled L2-egress bytes per traced ray
    by each launch's rays; the committed profile must carry them."""
    import json
    prof = json.load(open(os.path.join(ROOT, "profiles", "traversal_traffic.json")))
    tr = prof["trace"]
    assert prof["build"] and tr["kernel"] == "k_trace_bounce_fast"
    assert tr["rays_per_frame"] == tr["extend_rays_per_frame"] + tr["shadow_rays_per_frame"]
    per_ray = (tr["read_bytes"] + tr["write_bytes"]) / tr["rays_per_frame"]
    assert abs(per_ray - tr["l2_egress_bytes_per_ray"]) < 0.01 and 0 < per_ray < 4096
