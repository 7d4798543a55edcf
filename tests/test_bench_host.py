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
        cals = json.load(f)["calibration"]
    for cal in cals.values():  # one counter calibration per precision's runs
        assert abs(cal["fetch_factor"] - 2.0) < 1e-3 and cal["write_factor"] == 1.0


def test_kernel_symbols_match_rocprof_names():
    sys.path.insert(0, ROOT)
    import bench
    shape = (512, 512, 512)
    k = {"axis": 1, "inv": 1, "ld": 0, "st": 0, "variant": 0}
    assert bench.kernel_symbol(k, shape, True) == DOMINANT
    assert bench.kernel_symbol(dict(k, axis=0, ld=42), shape, False) == \
        "void qsm::k_group<double, 512, 2>(qsm::GroupArgs)"
    assert bench.kernel_symbol(dict(k, axis=2, inv=0, ld=26), shape, True) == \
        "void qsm::k_multi<float, 512, 26, 2>(qsm::MultiArgs)"
    assert bench.kernel_symbol(dict(k, axis=2, inv=0, ld=25), shape, True) == \
        "void qsm::k_multi<float, 512, 25, 0>(qsm::MultiArgs)"
    assert bench.kernel_symbol(dict(k, axis=0, inv=1, variant=4), shape, False) == \
        "void qsm::k_pass_pipe<double, 512, 0, true, 4>(qsm::PassArgs, int)"


def test_step_traffic_sums_launches_times_bytes():
    """Measured bytes per launch times launches per step, the algorithmic
    bytes standing in for a kernel without counters, over the step time."""
    sys.path.insert(0, ROOT)
    import bench
    table = [{"symbol": "a", "launches": 3.0, "ms": 1.0, "alg_bytes": 3e9},
             {"symbol": "b", "launches": 1.0, "ms": 0.5, "alg_bytes": 2e9}]
    t = bench.step_traffic(table, 2.0, {"a": 1.1e9})
    assert abs(t["measured_bytes_per_step"] - (3.3e9 + 2e9)) < 1
    assert t["kernels_without_counters"] == ["b"]
    assert abs(t["alg_GBs"] - 5e9 / 2e-3 / 1e9) < 1e-9
    assert abs(t["measured_frac"] - 5.3e9 / 2e-3 / 1e9 / 8000.0) < 1e-12
    assert t["kernel_ms_per_step"] == 1.5
