"""The C++ host driver (q-vti_amd/qvti_shots, host/qvti_shots.cpp): the
reference host's forward-modelling main on one GPU.  A 2-shot run over a raw
interior vp file is checked against the FP32 / FP64 oracle on the model and
geometry the reference host builds (edge-replicated vp, vs = vp/1.5, Qp 200,
Qs 150, rho 1800; x-force at s_iz + 10; receivers every dr-th interior cell)."""
import os
import subprocess

import numpy as np
import pytest

import oracle as O
import qvti
from cases import rel_l2

pytestmark = pytest.mark.gpu

HERE = os.path.dirname(os.path.abspath(__file__))
EXE = os.path.join(os.path.dirname(HERE), "q-vti_amd", "qvti_shots")


def test_host_driver_matches_the_oracle(tmp_path):
    nx, ny, nz, pml, nt, dr, dt, f0, d = 24, 12, 16, 4, 40, 2, 5e-4, 30.0, 10.0
    rng = np.random.default_rng(5)
    vp_int = (2600 + 600 * rng.random((nx, ny, nz))).astype(np.float32)
    vp_file = str(tmp_path / "vp.bin")
    qvti.write_raw(vp_file, vp_int)
    args = [EXE, "--nx", nx, "--ny", ny, "--nz", nz, "--pml", pml, "--nt", nt, "--dx", d, "--dy", d, "--dz", d,
            "--dt", dt, "--f0", f0, "--nsx", 2, "--dsx", 5, "--nsx0", 8, "--nsy0", 6, "--depth", 2, "--dr", dr,
            "--vp", vp_file, "--out", str(tmp_path)]
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    R = (nx // dr) * (ny // dr)
    recs = [[np.fromfile(str(tmp_path / ("%s_shot%d.bin" % (c, s))), np.float32).reshape(nt, R)
             for c in ("vx", "vy", "vz")] for s in range(2)]
    assert open(str(tmp_path / "source_xy.txt")).read().split() == ["8", "6", "13", "6"]

    ntx, nty, ntz = nx + 2 * pml, ny + 2 * pml, nz + 2 * pml
    vp = qvti.pad_model(vp_int, pml)
    model = dict(vp=vp, vs=(vp / np.float32(1.5)).astype(np.float32), rho=np.full(vp.shape, 1800, np.float32),
                 gp=qvti.gamma_from_q(np.full(vp.shape, 200, np.float32)),
                 gs=qvti.gamma_from_q(np.full(vp.shape, 150, np.float32)))
    rx, ry = qvti.reference_receivers(pml, nx // dr, ny // dr, dr)
    w0 = float(np.float32(2 * np.float32(3.1415926) * np.float32(f0)))
    for s, sx in enumerate((pml + 8, pml + 13)):
        shot = dict(nx=ntx, ny=nty, nz=ntz, pml=pml, nt=nt, dx=d, dy=d, dz=d, dt=dt, f0=f0, w0=w0,
                    src_kind=O.SRC_FX, sx=sx, sy=pml + 6, sz=pml + 2 + 10, rx=rx, ry=ry,
                    rz=np.full(R, pml + 2, np.int32), rik=qvti.ricker(nt, dt, f0))
        r32 = O.run(shot, model, "f32")["records"]
        r64 = O.run(shot, model, "f64")["records"]
        for c in range(3):
            floor = rel_l2(r32[c], r64[c])
            e32, e64 = rel_l2(recs[s][c], r32[c]), rel_l2(recs[s][c], r64[c])
            print("shot %d component %d: vs FP32 %.2e, vs FP64 %.2e (floor %.2e)" % (s, c, e32, e64, floor))
            assert e64 < max(1e-5, 2 * floor) and e32 < max(1e-5, 3 * floor), (s, c)


def test_host_driver_migrates_each_shot(tmp_path):
    """--rtm-type 1 (Q-compensated: both wavefields in the Q-compensating
    medium), cross-correlation, checkpoints every 7 steps and the direct wave
    removed with a homogeneous vp = 2600 run: the driver's image equals
    qvti_rtm called from Python on the driver's own records minus the direct
    wave, bit for bit, and the one-shot stack equals the shot's image."""
    nx, ny, nz, pml, nt, dr, dt, f0, d = 24, 12, 16, 4, 40, 2, 5e-4, 30.0, 10.0
    rng = np.random.default_rng(6)
    vp_int = (2600 + 600 * rng.random((nx, ny, nz))).astype(np.float32)
    vp_file = str(tmp_path / "vp.bin")
    qvti.write_raw(vp_file, vp_int)
    args = [EXE, "--nx", nx, "--ny", ny, "--nz", nz, "--pml", pml, "--nt", nt, "--dx", d, "--dy", d, "--dz", d,
            "--dt", dt, "--f0", f0, "--nsx", 1, "--nsx0", 8, "--nsy0", 6, "--depth", 2, "--dr", dr,
            "--vp", vp_file, "--out", str(tmp_path), "--rtm-type", 1, "--imaging-condition", 4, "--checkpoint", 7,
            "--direct-vp", 2600]
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    ntx, nty, ntz = nx + 2 * pml, ny + 2 * pml, nz + 2 * pml
    R = (nx // dr) * (ny // dr)
    recs = np.array([np.fromfile(str(tmp_path / ("%s_shot0.bin" % c)), np.float32).reshape(nt, R)
                     for c in ("vx", "vy", "vz")])
    image = np.fromfile(str(tmp_path / "image_shot0.bin"), np.float32).reshape(ntx, nty, ntz)
    illum = np.fromfile(str(tmp_path / "illum_shot0.bin"), np.float32).reshape(ntx, nty, ntz)
    stack = np.fromfile(str(tmp_path / "image_stack.bin"), np.float32).reshape(ntx, nty, ntz)

    vp = qvti.pad_model(vp_int, pml)
    gp = qvti.gamma_from_q(np.full(vp.shape, 200, np.float32))
    gs = qvti.gamma_from_q(np.full(vp.shape, 150, np.float32))
    rho = np.full(vp.shape, 1800, np.float32)
    rx, ry = qvti.reference_receivers(pml, nx // dr, ny // dr, dr)
    shot = qvti.make_shot(qvti.SRC_FX, (pml + 8, pml + 6, pml + 2 + 10), rx, ry, np.full(R, pml + 2, np.int32))
    w0 = float(np.float32(2 * np.float32(3.1415926) * np.float32(f0)))
    rik = qvti.ricker(nt, dt, f0)

    def engine(v, media):
        e = qvti.Engine(ntx, nty, ntz, pml, d, d, d, dt, f0, w0)
        e.set_media(media)
        e.set_model(v, gp, (v / np.float32(1.5)).astype(np.float32), gs, rho)
        return e

    with engine(np.full(vp.shape, 2600, np.float32), qvti.MEDIA_VISCO) as e:
        direct = e.forward(nt, rik, shot)
    with engine(vp, qvti.MEDIA_QCOMP) as es, engine(vp, qvti.MEDIA_QCOMP) as er:
        ref, ref_ill = qvti.rtm(es, er, nt, rik, shot, list(recs - direct), qvti.IC_XCORR, qvti.FIELD_PRESSURE, 7, 0.01)
    print("image max %.3e; norm of the direct-wave run over that of the records %.2f" % (
        np.abs(ref).max(), np.linalg.norm(direct) / np.linalg.norm(recs)))
    assert np.abs(ref).max() > 0
    assert np.array_equal(image, ref) and np.array_equal(illum, ref_ill) and np.array_equal(stack, image)


def test_host_driver_writes_interior_snapshots(tmp_path):
    """--snap-every 10 --snap-field 5: the driver writes pzz on the interior
    every 10 steps of each shot, [nsnap][nx][ny][nz], equal bit for bit to the
    engine's snapshot hook on the same shot run from Python."""
    nx, ny, nz, pml, nt, dr, dt, f0, d = 24, 12, 16, 4, 35, 2, 5e-4, 30.0, 10.0
    args = [EXE, "--nx", nx, "--ny", ny, "--nz", nz, "--pml", pml, "--nt", nt, "--dx", d, "--dy", d, "--dz", d,
            "--dt", dt, "--f0", f0, "--nsx", 1, "--nsx0", 8, "--nsy0", 6, "--depth", 2, "--dr", dr,
            "--vp-const", 2800, "--out", str(tmp_path), "--snap-every", 10, "--snap-field", 5]
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr)
    assert r.returncode == 0, r.stderr
    snaps = np.fromfile(str(tmp_path / "snap_pzz_shot0.bin"), np.float32).reshape(4, nx, ny, nz)
    ntx, nty, ntz = nx + 2 * pml, ny + 2 * pml, nz + 2 * pml
    vp = np.full((ntx, nty, ntz), 2800, np.float32)
    R = (nx // dr) * (ny // dr)
    rx, ry = qvti.reference_receivers(pml, nx // dr, ny // dr, dr)
    shot = qvti.make_shot(qvti.SRC_FX, (pml + 8, pml + 6, pml + 2 + 10), rx, ry, np.full(R, pml + 2, np.int32))
    w0 = float(np.float32(2 * np.float32(3.1415926) * np.float32(f0)))
    with qvti.Engine(ntx, nty, ntz, pml, d, d, d, dt, f0, w0) as e:
        e.set_model(vp, qvti.gamma_from_q(np.full(vp.shape, 200, np.float32)), (vp / np.float32(1.5)).astype(np.float32),
                    qvti.gamma_from_q(np.full(vp.shape, 150, np.float32)), np.full(vp.shape, 1800, np.float32))
        ref = e.set_snapshot(5, 10, nt)
        e.forward(nt, qvti.ricker(nt, dt, f0), shot)
    print("snapshot max |pzz| %.3e" % np.abs(ref).max())
    assert np.abs(ref[-1]).max() > 0 and np.array_equal(snaps, ref)


@pytest.mark.parametrize("media", [2, 3, 4])
def test_host_driver_models_in_the_reference_media(tmp_path, media):
    """--media 2|3|4 (the reference's media_type: dispersion-only, lossless,
    Q-compensating): the driver's records equal the engine's in that medium,
    run from Python on the same shot, bit for bit, and differ from the
    viscoelastic run's."""
    nx, ny, nz, pml, nt, dr, dt, f0, d = 24, 12, 16, 4, 30, 2, 5e-4, 30.0, 10.0
    args = [EXE, "--nx", nx, "--ny", ny, "--nz", nz, "--pml", pml, "--nt", nt, "--dx", d, "--dy", d, "--dz", d,
            "--dt", dt, "--f0", f0, "--nsx", 1, "--nsx0", 8, "--nsy0", 6, "--depth", 2, "--dr", dr,
            "--vp-const", 2800, "--qp", 30, "--qs", 20, "--out", str(tmp_path), "--media", media]
    r = subprocess.run([str(a) for a in args], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    R = (nx // dr) * (ny // dr)
    recs = np.array([np.fromfile(str(tmp_path / ("%s_shot0.bin" % c)), np.float32).reshape(nt, R)
                     for c in ("vx", "vy", "vz")])
    ntx, nty, ntz = nx + 2 * pml, ny + 2 * pml, nz + 2 * pml
    vp = np.full((ntx, nty, ntz), 2800, np.float32)
    rx, ry = qvti.reference_receivers(pml, nx // dr, ny // dr, dr)
    shot = qvti.make_shot(qvti.SRC_FX, (pml + 8, pml + 6, pml + 2 + 10), rx, ry, np.full(R, pml + 2, np.int32))
    w0 = float(np.float32(2 * np.float32(3.1415926) * np.float32(f0)))
    out = {}
    for m in (media - 1, qvti.MEDIA_VISCO):
        with qvti.Engine(ntx, nty, ntz, pml, d, d, d, dt, f0, w0) as e:
            e.set_media(m)
            e.set_model(vp, qvti.gamma_from_q(np.full(vp.shape, 30, np.float32)),
                        (vp / np.float32(1.5)).astype(np.float32),
                        qvti.gamma_from_q(np.full(vp.shape, 20, np.float32)), np.full(vp.shape, 1800, np.float32))
            out[m] = e.forward(nt, qvti.ricker(nt, dt, f0), shot)
    assert np.array_equal(recs, out[media - 1])
    assert not np.array_equal(recs, out[qvti.MEDIA_VISCO])
