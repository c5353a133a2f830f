"""CPU checks of bench.py: the algorithmic-bytes formula against the per-config
figures it is defined by, the workload table, and the one-GPU refusal."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
sys.path.insert(0, ROOT)

import bench  # noqa: E402


@pytest.mark.parametrize("dims,pml,vti,want", [
    ((64, 64, 64), 10, False, 996),
    ((128, 128, 128), 20, True, 1056),
    ((256, 256, 256), 20, False, 966),
    ((256, 256, 256), 20, True, 1026),
    ((400, 400, 200), 20, True, 1022),
    ((160, 160, 128), 20, True, 1048),
])
def test_b_alg_matches_the_per_config_figures(dims, pml, vti, want):
    """B_alg = 72 (9 R2C) + 432 (54 C2R) + 72 (dt terms) + 72 (v update)
    + 180 + 48 + 4C (stress update, C = 15 iso / 30 VTI) + 64 sum(2 pml / n),
    rounded, for each benchmark configuration."""
    assert round(bench.b_alg(*dims, pml, vti)) == want


def test_workloads_are_the_benchmark_configurations():
    assert bench.WORKLOADS["vti256"][:4] == (256, 256, 256, 20)
    assert bench.WORKLOADS["vti400"][:4] == (400, 400, 200, 20)
    assert bench.WORKLOADS["vti160"][:4] == (160, 160, 128, 20)
    assert bench.WORKLOADS["vti128"][:4] == (128, 128, 128, 20)
    assert bench.WORKLOADS["iso64"][:4] == (64, 64, 64, 10)
    for name, (nx, ny, nz, pml, d, dt, vti) in bench.WORKLOADS.items():
        assert dt == 1e-3 and vti == (name != "iso64")


def test_more_than_one_gpu_is_refused_before_any_device_work():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr and r.stdout == ""
