"""bench.py's command-line contract that runs without a GPU: more than one
GPU is refused with a message, and the roofline traffic is looked up for
the decomposition the autotuner picked (None when it was never measured)."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, cwd=ROOT, timeout=120)
    assert r.returncode != 0
    assert "one GPU" in (r.stdout + r.stderr)


def test_traffic_lookup():
    sys.path.insert(0, ROOT)
    import bench

    t = bench.traffic_bytes("synthetic", 400, 256)
    r, l = bench.TRAFFIC_KIB[("synthetic", "R", 400)], bench.TRAFFIC_KIB[("synthetic", "L", 256)]
    expect = 0.5 * ((r[0] + l[0]) * bench.FETCH_SCALE + (r[1] + l[1]) * bench.WRITE_SCALE) * 1024
    assert t == round(expect)
    assert bench.traffic_bytes("synthetic", 401, 256) is None
