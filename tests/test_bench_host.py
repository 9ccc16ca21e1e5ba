"""The benchmark's host logic on the CPU: method names map to the solver set-up
they name, --gpus above 1 is refused before anything touches a GPU, and the
roofline's kernel and HBM traffic are read from the committed rocprofv3
summaries."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


@pytest.mark.parametrize("name,want", [
    ("fp64", dict(mode="fp64", solver="bicgstab", preconditioner="ilu")),
    ("mixed-cg", dict(mode="mixed", solver="cg", preconditioner="ilu")),
    ("mixed-cg-mg", dict(mode="mixed", solver="cg", preconditioner="multigrid")),
    ("fp64-gmres-mg", dict(mode="fp64", solver="gmres", preconditioner="multigrid")),
    ("mixed-gmres-mg", dict(mode="mixed", solver="gmres", preconditioner="multigrid")),
    ("learned-cg-mg", dict(mode="learned", solver="cg", preconditioner="multigrid", fallbackRatio=0.5)),
])
def test_method_names(name, want):
    assert bench.solver_kw(name) == want


def test_setup_pass_through():
    class Recorder:
        def __init__(self):
            self.got = []

        def setup(self, k, v):
            self.got.append((k, v))

    class Args:
        setup = "multigridJacobiWeight=0.8,gmresRestart=40"

    r = Recorder()
    bench.apply_setup(r, Args())
    assert r.got == [("multigridJacobiWeight", "0.8"), ("gmresRestart", "40")]


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120, cwd=ROOT)
    assert p.returncode == 2
    assert "one GPU" in p.stderr and p.stdout == ""


def test_roofline_kernel_and_traffic_from_committed_profiles():
    assert bench.dominant_from_trace() == "phase::kMlResid<float, double>"
    assert bench.TRACE_KEYS["phase::kMlResid<float, double>"] == "mlresid32"
    traffic = bench.measured_traffic("phase::kMlResid<float, double>")
    # HBM bytes per launch from the FETCH_SIZE / WRITE_SIZE passes: within
    # 0.8-1.3 of the kernel's 60,037,928 algorithmic bytes
    assert 0.8 * 60037928 < traffic < 1.3 * 60037928


def test_metric_is_the_baseline_metric():
    import json
    with open(os.path.join(ROOT, "BASELINE.json")) as f:
        assert bench.METRIC == json.load(f)["metric"]
