"""The bench script's host-side contract (no GPU): one GPU only, and the
roofline's measured traffic comes from the committed counter summary."""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, REPO)


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(REPO, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


@pytest.mark.parametrize("kernel", ["k_ps_bulk", "k_ps_fused"])
@pytest.mark.parametrize("storage", ["f32", "f16"])
def test_measured_traffic_is_read_from_the_counter_summary(kernel, storage):
    import bench

    t = bench.measured_traffic(kernel, storage)
    assert t is not None
    # within a few percent of the algorithmic bytes of the sphere's update
    alg = (4161536 if kernel == "k_ps_bulk" else 4194304) * (216 if storage == "f32" else 108)
    assert 1.0 <= t / alg < 1.2
    assert bench.measured_traffic("no_such_kernel", storage) is None
