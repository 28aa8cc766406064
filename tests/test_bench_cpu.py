"""bench.py's contract pieces that need no GPU: the refusal of several GPUs
and the traffic figures it reads from the committed counter summary."""
import json
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
sys.path.insert(0, str(ROOT))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120, cwd=ROOT)
    assert r.returncode == 2
    msg = json.loads(r.stdout.strip().splitlines()[-1])
    assert "one GPU" in msg["error"]


def test_traffic_per_step_from_counter_summary():
    """The counter summary gives the collide-stream traffic per step for the
    profiled tile shape; another shape gets none (no stale figure)."""
    import bench
    for prec, bytes_per_lup in (("fp16", 2 * 19 * 2), ("fp32", 2 * 19 * 4)):
        t = bench.measured_traffic(prec, 256, 2)
        assert t is not None, f"no {prec} traffic line in the counter summary"
        algorithmic = bytes_per_lup * 512 * 256 * 128
        assert 0.9 < t / algorithmic < 1.1
        assert bench.measured_traffic(prec, 64, 1) is None


def test_reference_deck_is_the_workload_relabelled():
    """The CPU baseline's deck is the benchmark channel with y and z swapped:
    same sizes, walls, sphere and distance from the wall."""
    import bench
    w, r = bench.WORKLOAD, bench.REFERENCE_DECK
    assert (r["nx"], r["ny"], r["nz"]) == (w["nx"], w["nz"], w["ny"])
    assert (r["walls_y"], r["walls_z"]) == (w["walls_z"], w["walls_y"])
    assert r["y0"] == w["z0"] and r["angle"] == w["angle"] and "plane_xz" not in r
    for k in ("radius", "b1", "beta", "tau", "eps", "hc", "hmin", "planar"):
        assert r[k] == w[k], k


def test_ensemble_traffic_per_shape_from_counter_summary():
    import bench
    nodes = 200219976  # the 64 runs of seed 0
    # every shape the batch tuner tries has a counter run
    for w in (64, 128, 256):
        for x in (0, 1, 2):
            t = bench.measured_ensemble_traffic(w, x)
            assert t is not None, (w, x)
            assert 0.95 < t / (76 * nodes) < 1.3
    for shape in ((0, 0), (0, 1)):
        assert bench.measured_ensemble_traffic(*shape) is not None
    assert bench.measured_ensemble_traffic(512, 1) is None  # no counter run for that shape
