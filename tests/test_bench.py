"""bench.py's contract pieces that run without a GPU: the one-GPU refusal, the
committed HBM-traffic entry for the headline kernel as built today, and the
CPU baseline's report."""
import json
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert p.returncode == 2
    line = json.loads(p.stdout.strip().splitlines()[-1])
    assert "one GPU" in line["error"]


def test_traffic_summary_matches_the_default_blocked_kernel():
    """The headline's `roofline.traffic` is looked up by a kernel id made of the
    blocked loop's tile; the committed counter summary must hold that id."""
    import bench
    import xgrid_amd
    from xgrid_amd import workloads as W
    from xgrid_amd.backend import lower_kernel
    xgrid_amd.init(precision="double", cacheroot=os.path.join(ROOT, ".xgrid"))
    low, _ = lower_kernel(W.make_cavity())
    L = low.loops[0]
    n = 2049
    kid = f"{L.kernel}:{L.core[0]}x{L.core[1]}:h{L.halo}:t{L.tmax}:w{L.threads}:n{n}"
    t = bench.measured_traffic(kid)
    assert t is not None, f"profiles/traffic.json has no entry for {kid}"
    entry = json.load(open(os.path.join(ROOT, "profiles", "traffic.json")))["kernels"][kid]
    # HBM bytes of a launch: at least the level read and written once, far below the algorithmic bytes
    assert 2 * 2049 * 2049 * 8 < t < entry["algorithmic_bytes"]
    assert bench.measured_traffic("no-such-kernel") is None


def test_cpu_baseline_reports_cores_and_sample():
    import bench
    import xgrid_amd
    xgrid_amd.init(precision="double", cacheroot=os.path.join(ROOT, ".xgrid"))
    r = bench.cpu_baseline(65, 2e-6, 0.2)
    assert r["kind"] == "port" and r["unit"] == "cell-updates/s" and r["value"] > 0
    assert r["cores"] >= 1 and "65x65" in r["sample"]
    assert r["single_thread"]["cores"] == 1 and r["single_thread"]["value"] > 0


def test_roofline_entry_reports_measured_traffic_and_valu_beside_algorithmic_rate():
    """A blocked kernel's algorithmic rate may exceed the HBM peak; the entry then
    carries the measured HBM rate (below the peak) and the VALU figure."""
    import bench
    kid = "xg_tb0:48x40:h12:t10:w1024:n2049"
    r = bench.roofline_entry(kid, "xg_tb0", 1006796960.0, 0.111)
    assert r["bound"] == "hbm" and r["unit"] == "GB/s" and r["peak"] == 8000.0
    assert abs(r["frac"] - r["achieved"] / 8000.0) < 1e-12 and r["frac"] > 1.0
    assert 0 < r["hbm_measured_frac"] < 0.5 and r["traffic"] < 1006796960.0
    assert 0 < r["valu_busy"] < 1 and "note" in r
