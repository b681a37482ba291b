"""bench.py's contract pieces that run without a GPU: the multi-GPU refusal
and the link between its dominant-kernel names and the committed rocprofv3
summaries it reads the measured traffic from."""
import subprocess
import sys

from conftest import ROOT


def test_bench_refuses_more_than_one_gpu():
    r = subprocess.run([sys.executable, str(ROOT / "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr


def test_bench_kernel_names_are_in_the_committed_profiles():
    sys.path.insert(0, str(ROOT))
    import bench
    traffic = bench.TRAFFIC_FILE.read_text()
    stats = (ROOT / "profiles" / "kernel_stats_inner_iteration_256x192x256.md").read_text()
    for stage, kernel in bench.STAGE_KERNEL.items():
        if stage == "rhs":
            continue  # the device-resident iteration forms the RHS inside the sweep pre-pass
        assert f"`{kernel}`" in stats, kernel
        assert bench.measured_traffic(kernel), kernel
        assert f"`{kernel}`" in traffic, kernel
        assert bench.valu_floor_us(kernel), kernel
