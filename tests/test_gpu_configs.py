"""GPU parity on the benchmark configurations against committed fixtures.

Isotropic 64^3 configuration (homogeneous vp 3000, vs 1730, rho 2200, Qp 100,
Qs 50, f0 25 Hz, dx 10 m, dt 1 ms, 500 steps, explosive Ricker at the centre,
10-cell C-PML, 64 receivers along x at y = 32, z = 5 recording vz): the
receiver gathers must be within 1e-5 relative L2 of the FP32 and the FP64
oracle (tests/golden/cfg1_records.npz, written by make_cfg1_golden.py).
"""
import os

import numpy as np
import pytest

import qvti
from cases import rel_l2

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
PI = 3.1415926


def test_config1_iso64_500_steps_vs_golden():
    g = np.load(os.path.join(HERE, "golden", "cfg1_records.npz"))
    n, pml, f0, nt = 64, 10, 25.0, 500
    m = qvti.make_model(0, 0, (n, n, n), [3000, 1730, 2200, 100, 50])
    rik = qvti.ricker(nt, 1e-3, f0)
    assert np.array_equal(rik, g["rik"])
    with qvti.Engine(n, n, n, pml, 10.0, 10.0, 10.0, 1e-3, f0, float(np.float32(2 * PI * f0))) as e:
        e.set_model_dict(m)
        rec = e.forward(nt, rik, qvti.make_shot(qvti.SRC_EXPLOSIVE, (32, 32, 32), np.arange(64), np.full(64, 32),
                                                np.full(64, 5)))
    e32 = rel_l2(rec[2], g["rec_f32"][2])
    e64 = rel_l2(rec[2], g["vz_f64"])
    floor = rel_l2(g["rec_f32"][2], g["vz_f64"])
    print("config 1 vz: vs FP32 oracle %.3e, vs FP64 oracle %.3e (FP32 oracle vs FP64 %.3e)" % (e32, e64, floor))
    assert e32 < 1e-5 and e64 < 1e-5
