"""GPU tests of the propagation-medium option and reverse-time migration
(qvti_set_media, qvti_rtm), against the oracle and against the kinematics of
a flat reflector.

The reference only declares its RTM mode (RTM_type, imaging_condition,
media_type, reconstruction flags at viscoelas_3D.cpp:42-65; checkpoints
every 500 steps at viscoelas_3D.cpp:223-229).  The oracle computes the same
migration from whole-grid wavefield histories: the source run S(t) and the
receiver run R(tau) whose sources are the records injected time-reversed at
the receiver cells (after the stress update, entering the next step's
V - V_prev like the shot's own source), imaged at equal physical time
t = nt - 1 - tau.
"""
import numpy as np
import pytest

import oracle as O
import qvti
from cases import rel_l2, small_case

pytestmark = pytest.mark.gpu
TOL = 1e-5


def _engine(shot, model, media=0):
    e = qvti.Engine(shot["nx"], shot["ny"], shot["nz"], shot["pml"], shot["dx"], shot["dy"], shot["dz"],
                    shot["dt"], shot["f0"], shot["w0"], vti="eps" in model)
    e.set_media(media)
    e.set_model_dict(model)
    return e


def _shot(shot):
    return qvti.make_shot(shot["src_kind"], (shot["sx"], shot["sy"], shot["sz"]), shot["rx"], shot["ry"], shot["rz"])


@pytest.mark.parametrize("media", [qvti.MEDIA_DISPERSION, qvti.MEDIA_LOSSLESS, qvti.MEDIA_QCOMP])
def test_media_match_the_oracle(media):
    """Dispersion-only, lossless and Q-compensating media (the reference's
    media_type 2, 3, 4) against the oracle in the same medium."""
    shot, model = small_case(20, 24, 32, 4, 30, O.SRC_EXPLOSIVE, True)
    ref = O.run(shot, model, "f64", state=True, media=media)
    base = O.run(shot, model, "f64", state=True)
    with _engine(shot, model, media) as e:
        e.begin(shot["nt"], shot["rik"], [_shot(shot)])
        e.step(shot["nt"])
        rec = e.finish()[0]
        st = e.state()
    err, diff = rel_l2(st, ref["state"]), rel_l2(ref["state"], base["state"])
    print("media %d: state vs FP64 oracle %.2e (differs from the constant-Q scheme by %.2e)" % (media, err, diff))
    assert err < TOL and rel_l2(rec, ref["records"]) < TOL
    assert diff > 100 * TOL  # the medium really changed the run


def _oracle_rtm(shot, model, recs, field, condition, eps, media_r=0):
    """Image from the oracle's whole-grid histories (FP64)."""
    nt = shot["nt"]
    S = O.run(shot, model, "f64", hist_field=field)["hist"]
    rshot = dict(shot, rik=np.zeros(nt, np.float32), rx=shot["rx"][:1], ry=shot["ry"][:1], rz=shot["rz"][:1])
    inj = dict(x=shot["rx"], y=shot["ry"], z=shot["rz"])
    for c, name in enumerate(("vx", "vy", "vz")):
        inj[name] = None if recs[c] is None else np.ascontiguousarray(recs[c][::-1])
    R = O.run(rshot, model, "f64", hist_field=field, inject=inj, media=media_r)["hist"][::-1]  # R at physical time t
    if condition == qvti.IC_EXA:
        a = np.abs(S)
        tex = np.argmax(a, axis=0)
        smax = np.take_along_axis(S, tex[None], 0)[0]
        thr = eps * np.abs(smax).max()
        on = np.abs(smax) > thr
        img = np.where(on, np.take_along_axis(R, tex[None], 0)[0] / np.where(on, smax, 1), 0)
        # left out of the comparison: cells whose two largest |S| nearly tie
        # (the excitation time is not determined by the scheme) and cells at
        # the imaging threshold
        top2 = np.sort(a, axis=0)[-2:]
        tie = (top2[1] - top2[0] < 1e-3 * top2[1]) | (np.abs(np.abs(smax) - thr) < 1e-2 * thr)
        return img, smax, tie
    img = (S * R).sum(0)
    ill = (S * S).sum(0)
    if condition == qvti.IC_SRCNORM:
        img = img / (ill + eps * ill.max())
    return img, ill, None


@pytest.mark.parametrize("condition,field,comps,ckpt,media_r", [
    (qvti.IC_XCORR, qvti.FIELD_PRESSURE, (2,), 16, 0),
    (qvti.IC_XCORR, 2, (0, 1, 2), 13, 0),
    (qvti.IC_SRCNORM, qvti.FIELD_PRESSURE, (2,), 25, 0),
    (qvti.IC_XCORR, qvti.FIELD_PRESSURE, (2,), 60, qvti.MEDIA_QCOMP),
    (qvti.IC_EXA, qvti.FIELD_PRESSURE, (2,), 16, 0),
])
def test_rtm_matches_the_oracle(condition, field, comps, ckpt, media_r):
    """Checkpointed source reconstruction (segments of `ckpt` steps, the last
    one partial), time-reversed injection of the recorded components and the
    imaging condition, against the oracle's image from full histories."""
    shot, model = small_case(24, 30, 32, 4, 60, O.SRC_EXPLOSIVE, True)
    nt = shot["nt"]
    with _engine(shot, model) as e:
        recs = e.forward(nt, shot["rik"], _shot(shot))
    recs = [recs[c] if c in comps else None for c in range(3)]
    print("case: condition %d field %d components %s checkpoint %d receiver medium %d" % (
        condition, field, comps, ckpt, media_r), flush=True)
    with _engine(shot, model) as es, _engine(shot, model, media_r) as er:
        img, ill = qvti.rtm(es, er, nt, shot["rik"], _shot(shot), recs, condition, field, ckpt, 0.01)
    ref, ref_ill, tie = _oracle_rtm(shot, model, recs, field, condition, 0.01, media_r)
    if tie is not None:
        keep = ~tie
        print("EXA: %d of %d cells with near-tied excitation times left out" % (tie.sum(), tie.size))
        err, e_ill = rel_l2(img[keep], ref[keep]), rel_l2(ill[keep], ref_ill[keep])
    else:
        err, e_ill = rel_l2(img, ref), rel_l2(ill, ref_ill)
    print("image vs FP64 oracle %.2e, illumination %.2e" % (err, e_ill))
    assert np.isfinite(img).all() and np.abs(ref).max() > 0
    assert err < TOL and e_ill < TOL


def test_rtm_images_a_flat_reflector_at_its_depth():
    """Two layers (vp 2500 / 3500 m/s across z = 40, lossless), explosive
    source and a receiver line near the surface, the direct wave removed from
    the data; migration in the upper layer's velocity.  The cross-correlation
    image of the pressure under the source peaks at the interface."""
    nx, ny, nz, pml, d, dt, nt, f0 = 64, 32, 64, 10, 10.0, 1e-3, 360, 25.0
    zr, sz = 40, pml + 2
    shape = (nx, ny, nz)
    vp = np.full(shape, 2500.0, np.float32)
    vp[:, :, zr:] = 3500.0
    zero = np.zeros(shape, np.float32)
    true = dict(vp=vp, vs=(vp / 1.8).astype(np.float32), rho=np.full(shape, 2000, np.float32), gp=zero, gs=zero)
    flat = dict(true, vp=np.full(shape, 2500.0, np.float32), vs=np.full(shape, 2500.0 / 1.8, np.float32))
    sx, sy = nx // 2, ny // 2
    rx = np.arange(pml, nx - pml)
    shot = dict(nx=nx, ny=ny, nz=nz, pml=pml, nt=nt, dx=d, dy=d, dz=d, dt=dt, f0=f0,
                w0=float(np.float32(2 * 3.1415926 * f0)), src_kind=O.SRC_EXPLOSIVE, sx=sx, sy=sy, sz=sz,
                rx=rx, ry=np.full(len(rx), sy), rz=np.full(len(rx), sz), rik=qvti.ricker(nt, dt, f0))
    with _engine(shot, true) as e:
        data = e.forward(nt, shot["rik"], _shot(shot))
    with _engine(shot, flat) as e:
        direct = e.forward(nt, shot["rik"], _shot(shot))
    refl = data - direct
    with _engine(shot, flat) as es, _engine(shot, flat) as er:
        img, ill = qvti.rtm(es, er, nt, shot["rik"], _shot(shot), [None, None, refl[2]], qvti.IC_XCORR,
                            qvti.FIELD_PRESSURE, 90, 0.0)
    col = np.abs(img[sx, sy, sz + 6:nz - pml])
    z_peak = sz + 6 + int(np.argmax(col))
    print("image peak under the source at z = %d (interface between z = %d and %d); peak / median %.1f" % (
        z_peak, zr - 1, zr, col.max() / np.median(col)))
    assert abs(z_peak - (zr - 0.5)) <= 2.0
    assert col.max() > 5 * np.median(col)


def _rtm_case():
    shot, model = small_case(20, 24, 30, 4, 40, O.SRC_EXPLOSIVE, True)
    with _engine(shot, model) as e:
        recs = e.forward(shot["nt"], shot["rik"], _shot(shot))
    return shot, model, recs


def test_rtm_segments_do_not_change_the_image():
    """The source wavefield recomputed from checkpoints in segments of 7 steps
    (the last one partial) is the uninterrupted run bit for bit, so the
    cross-correlation image and illumination equal those of one segment
    covering all nt steps, bitwise."""
    shot, model, recs = _rtm_case()
    out = []
    for ckpt in (shot["nt"], 7):
        with _engine(shot, model) as es, _engine(shot, model) as er:
            out.append(qvti.rtm(es, er, shot["nt"], shot["rik"], _shot(shot), [None, None, recs[2]], qvti.IC_XCORR,
                                qvti.FIELD_PRESSURE, ckpt, 0.0))
    assert np.abs(out[0][0]).max() > 0
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


def test_rtm_image_is_linear_in_the_records():
    """The receiver wavefield is linear in the injected records and scaling
    by 2 is exact in binary floating point: doubling every record doubles the
    cross-correlation image bit for bit and leaves the illumination alone."""
    shot, model, recs = _rtm_case()
    out = []
    for scale in (1.0, 2.0):
        with _engine(shot, model) as es, _engine(shot, model) as er:
            out.append(qvti.rtm(es, er, shot["nt"], shot["rik"], _shot(shot), [scale * r for r in recs],
                                qvti.IC_XCORR, 2, 9, 0.0))
    assert np.array_equal(out[1][0], 2 * out[0][0]) and np.array_equal(out[1][1], out[0][1])


def test_rtm_rejects_bad_arguments():
    """One context for both wavefields, contexts of different grids, the
    reference's partial-correlation condition (2, not built), a checkpoint
    interval below 1, an unknown field and no records at all are refused
    before any device work."""
    shot, model, recs = _rtm_case()
    nt, rik, s = shot["nt"], shot["rik"], _shot(shot)
    other = dict(shot, nz=shot["nz"] + 2)
    m2 = {k: np.pad(v, ((0, 0), (0, 0), (0, 2)), mode="edge") for k, v in model.items()}
    with _engine(shot, model) as es, _engine(shot, model) as er, _engine(other, m2) as eo:
        bad = [
            lambda: qvti.rtm(es, es, nt, rik, s, recs, qvti.IC_XCORR, qvti.FIELD_PRESSURE, 10, 0.0),
            lambda: qvti.rtm(es, eo, nt, rik, s, recs, qvti.IC_XCORR, qvti.FIELD_PRESSURE, 10, 0.0),
            lambda: qvti.rtm(es, er, nt, rik, s, recs, 2, qvti.FIELD_PRESSURE, 10, 0.0),
            lambda: qvti.rtm(es, er, nt, rik, s, recs, qvti.IC_XCORR, qvti.FIELD_PRESSURE, 0, 0.0),
            lambda: qvti.rtm(es, er, nt, rik, s, recs, qvti.IC_XCORR, 10, 10, 0.0),
            lambda: qvti.rtm(es, er, nt, rik, s, [None, None, None], qvti.IC_XCORR, qvti.FIELD_PRESSURE, 10, 0.0),
        ]
        for k, call in enumerate(bad):
            with pytest.raises(qvti.QvtiError) as ex:
                call()
            assert ex.value.code == qvti.E_ARG, (k, ex.value)
        # the contexts stay usable
        img, _ = qvti.rtm(es, er, nt, rik, s, recs, qvti.IC_XCORR, qvti.FIELD_PRESSURE, 10, 0.0)
        assert np.isfinite(img).all() and np.abs(img).max() > 0
