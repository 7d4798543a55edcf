"""Host-side behaviour of bench.py that needs no GPU: the single-GPU guard and
the per-kernel traffic lookup the roofline object carries."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DOMINANT = "void qsm::k_pass<float, 512, 1, true, 0, 0, 0>(qsm::PassArgs)"


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU only" in r.stderr
    assert r.stdout == ""


def test_traffic_of_dominant_kernel_is_its_field_read_and_write():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.measured_traffic(DOMINANT)
    # one FP32 512^3 field read and written: 2 x 512 MiB, within 0.1%
    assert t is not None and abs(t / 2**30 - 1.0) < 1e-3, t
    with open(os.path.join(ROOT, "profiles", "traffic.json")) as f:
        cal = json.load(f)["calibration"]
    assert abs(cal["fetch_factor"] - 2.0) < 1e-3 and cal["write_factor"] == 1.0
