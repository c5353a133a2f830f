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

import oracle as O
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


def _run_window(g, m, shot_kind, src, rx, ry, rz, nt, kspace=None, state=False):
    """GPU records and FP32 / FP64 oracle records of the first nt steps (with
    state: the three wavefield states [9][nx][ny][nz] follow)."""
    rik = qvti.ricker(nt, g["dt"], g["f0"])
    with qvti.Engine(g["nx"], g["ny"], g["nz"], g["pml"], g["dx"], g["dy"], g["dz"], g["dt"], g["f0"], g["w0"],
                     vti=g["vti"]) as e:
        e.set_model_dict(m)
        if kspace:
            e.set_kspace(kspace["vref"])
        rec = e.forward(nt, rik, qvti.make_shot(shot_kind, src, rx, ry, rz))
        st = e.state() if state else None
    shot = dict(nx=g["nx"], ny=g["ny"], nz=g["nz"], pml=g["pml"], nt=nt, dx=g["dx"], dy=g["dy"], dz=g["dz"],
                dt=g["dt"], f0=g["f0"], w0=g["w0"], src_kind=shot_kind, sx=src[0], sy=src[1], sz=src[2],
                rx=rx, ry=ry, rz=rz, rik=rik)
    model = {k: m[k] for k in ("vp", "gp", "vs", "gs", "rho")}
    if g["vti"]:
        model["eps"], model["delta"] = m["eps"], m["delta"]
    o32 = O.run(shot, model, "f32", kspace=kspace, state=state)
    o64 = O.run(shot, model, "f64", kspace=kspace, state=state)
    if state:
        return rec, o32["records"], o64["records"], st, o32["state"], o64["state"]
    return rec, o32["records"], o64["records"]


def test_config2_vti128_layered_vs_golden():
    """VTI 128^3, 6 seeded flat layers (vp U[1800,4500], vs = vp/1.73, Gardner
    rho, eps U[0,0.25], delta U[0,eps], Qp U[30,200], Qs = Qp/2), dx 10 m,
    20-cell C-PML, Ricker 20 Hz vertical force at (64,64,10), 128x128
    receivers at z = 20, 200 steps of 0.5 ms (1 ms exceeds the scheme's
    stability limit for these speeds), against the committed FP32/FP64 oracle
    gathers (tests/golden/cfg2_records.npz, make_cfg2_golden.py; vz at every
    4th receiver).  Far receivers the wave reaches last carry mostly the FP32
    rounding floor, hence the floor-aware bound."""
    gold = np.load(os.path.join(HERE, "golden", "cfg2_records.npz"))
    n, pml, f0, nt, dt = 128, 20, 20.0, 200, 5e-4
    m = qvti.make_model(1, 11, (n, n, n), [1800, 4500, 0.25, 30, 200], nlay=6, vti=True)
    assert m["vp"].astype(np.float64).sum() == gold["vp_sum"]
    rik = qvti.ricker(nt, dt, f0)
    assert np.array_equal(rik, gold["rik"])
    ix, iy = np.meshgrid(np.arange(n), np.arange(n), indexing="ij")
    keep = ((ix % 4 == 0) & (iy % 4 == 0)).ravel()
    with qvti.Engine(n, n, n, pml, 10.0, 10.0, 10.0, dt, f0, float(np.float32(2 * PI * f0)), vti=True) as e:
        e.set_model_dict(m)
        rec = e.forward(nt, rik, qvti.make_shot(qvti.SRC_FZ, (64, 64, 10), ix.ravel(), iy.ravel(),
                                                np.full(n * n, pml)))
    vz = rec[2][:, keep]
    e32, e64 = rel_l2(vz, gold["vz_f32"]), rel_l2(vz, gold["vz_f64"])
    floor = rel_l2(gold["vz_f32"], gold["vz_f64"])
    print("config 2 vz: vs FP32 oracle %.3e, vs FP64 oracle %.3e (FP32 oracle vs FP64 %.3e)" % (e32, e64, floor))
    assert e64 < max(1e-5, 2 * floor) and e32 < max(1e-5, 3 * floor)


@pytest.mark.parametrize("workload,nt", [("vti256", 4), ("vti400", 3), ("vti160", 6)])
def test_config3_full_size_window(workload, nt):
    """The 256^3 and 400x400x200 single-shot configurations and the batched
    configuration's 160x160x128 grid exactly as benchmarked (seeded smooth
    random VTI model, Q in [20,150], eps/delta co-varying with vp, 20-cell
    C-PML, dx 12.5 m, dt 1 ms): the whole wavefield state (9 fields, every
    cell) after the first steps against the oracle.  A receiver line through
    the source would hold little but the injected wavelet this early; the
    state carries every spectral derivative of the step.  The records over
    200/160 steps are checked against fixtures in test_gpu_fullsize.py."""
    import bench

    g, m, _ = bench.make_workload(workload)
    src = (g["nx"] // 2, g["ny"] // 2, g["pml"] + 1)
    rx = np.arange(src[0] - 6, src[0] + 7)
    rec, r32, r64, st, s32, s64 = _run_window(g, m, qvti.SRC_FZ, src, rx, np.full(len(rx), src[1]),
                                              np.full(len(rx), src[2]), nt, state=True)
    e32, e64 = rel_l2(st, s32), rel_l2(st, s64)
    print("%s %d steps, whole state: vs FP32 oracle %.3e, vs FP64 oracle %.3e (FP32 oracle vs FP64 %.3e); "
          "records vs FP64 %.3e" % (workload, nt, e32, e64, rel_l2(s32, s64), rel_l2(rec, r64)))
    assert e32 < 1e-5 and e64 < 1e-5
    assert rel_l2(rec, r32) < 1e-5 and rel_l2(rec, r64) < 1e-5


def test_config3_kspace_window_at_the_specified_dt():
    """The 256^3 configuration at its specified 1 ms step with the opt-in
    k-space correction (vref = 1.25 x the fastest horizontal qP speed, which
    covers the Q dispersion of Q >= 20): 8 full-size steps against the FP32 /
    FP64 oracle in the same mode, on a receiver line through the source."""
    import bench

    g, m, _ = bench.make_workload("vti256")
    vh = float((m["vp"].astype(np.float64) * np.sqrt(1 + 2 * m["eps"].astype(np.float64))).max())
    src = (g["nx"] // 2, g["ny"] // 2, g["pml"] + 1)
    rx = np.arange(src[0] - 8, src[0] + 9)
    rec, r32, r64 = _run_window(g, m, qvti.SRC_FZ, src, rx, np.full(len(rx), src[1]), np.full(len(rx), src[2]), 8,
                                kspace=dict(vref=float(np.float32(1.25 * vh))))
    e32, e64 = rel_l2(rec, r32), rel_l2(rec, r64)
    print("vti256 k-space 8 steps: vs FP32 oracle %.3e, vs FP64 oracle %.3e" % (e32, e64))
    assert e32 < 1e-5 and e64 < 1e-5
