"""Writes tests/golden/cfg1_records.npz: receiver gathers of the isotropic
64^3 configuration (homogeneous vp 3000, vs 1730, rho 2200, Qp 100, Qs 50,
f0 = 25 Hz, dx = 10 m, dt = 1 ms, 500 steps, explosive Ricker at the grid
centre, 10-cell C-PML, 64 receivers x = 0..63 at y = 32, z = 5) from the FP32
and FP64 CPU oracle.  Run once in the build container:
    python tests/golden/make_cfg1_golden.py
"""
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(__file__))))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "q-vti_amd")]
import oracle as O  # noqa: E402
import qvti  # noqa: E402

PI = 3.1415926


def cfg1(nt=500):
    n, pml, f0 = 64, 10, 25.0
    model = qvti.make_model(0, 0, (n, n, n), [3000, 1730, 2200, 100, 50])
    shot = dict(nx=n, ny=n, nz=n, pml=pml, nt=nt, dx=10.0, dy=10.0, dz=10.0, dt=1e-3, f0=f0,
                w0=float(np.float32(2 * PI * f0)), src_kind=O.SRC_EXPLOSIVE, sx=32, sy=32, sz=32,
                rx=np.arange(64), ry=np.full(64, 32), rz=np.full(64, 5))
    shot["rik"] = qvti.ricker(nt, shot["dt"], f0)
    return shot, model


if __name__ == "__main__":
    shot, model = cfg1()
    t = time.time()
    r32 = O.run(shot, model, "f32")["records"]
    r64 = O.run(shot, model, "f64")["records"]
    np.savez_compressed(os.path.join(os.path.dirname(os.path.abspath(__file__)), "cfg1_records.npz"),
                        rec_f32=r32.astype(np.float32), vz_f64=r64[2].astype(np.float64),
                        rik=shot["rik"])
    print("done in %.0fs" % (time.time() - t))
