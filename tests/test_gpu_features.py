"""GPU tests of the engine's other entry points and hooks, against the oracle:
snapshots, the blow-up probe, the reference-compatible entry point
(cuda_forward_acoustic_3D's argument list), and exact linearity of the whole
path in the source at a full benchmark size."""
import ctypes

import numpy as np
import pytest

import oracle as O
import qvti
from cases import rel_l2, small_case

pytestmark = pytest.mark.gpu

TOL = 1e-5


def engine_for(shot, model, max_shots=1):
    e = qvti.Engine(shot["nx"], shot["ny"], shot["nz"], shot["pml"], shot["dx"], shot["dy"], shot["dz"],
                    shot["dt"], shot["f0"], shot["w0"], vti="eps" in model, max_shots=max_shots)
    e.set_model_dict(model)
    return e


def shot_of(shot):
    return qvti.make_shot(shot["src_kind"], (shot["sx"], shot["sy"], shot["sz"]), shot["rx"], shot["ry"], shot["rz"])


@pytest.mark.parametrize("field,every", [(2, 7), (3, 5), (7, 9)])
def test_snapshots_match_the_oracle(field, every):
    """qvti_set_snapshot: interior window [pml, n-pml)^3 of one field every
    `every` steps (it % every == 0, after the source), ix slowest - the
    reference's disabled dump (cuviscoelas_3D.cu:1615-1627)."""
    shot, model = small_case(24, 20, 32, 4, 30, O.SRC_EXPLOSIVE, True)
    ref = O.run(shot, model, "f64", snap_every=every, snap_field=field)
    with engine_for(shot, model) as e:
        snaps = e.set_snapshot(field, every, shot["nt"])
        e.begin(shot["nt"], shot["rik"], [shot_of(shot)])
        e.step(shot["nt"])
        e.finish()
    want = ref["snaps"].reshape(snaps.shape)
    assert snaps.shape[0] == (shot["nt"] + every - 1) // every
    for k in range(1, snaps.shape[0]):  # snapshot 0 is the first step's field (tiny)
        assert rel_l2(snaps[k], want[k]) < TOL, (k, rel_l2(snaps[k], want[k]))


def test_blow_up_probe_reports_an_unstable_run():
    """A time step far beyond the pseudospectral stability limit grows without
    bound; the probe (max |v| every `every` steps) turns that into
    QVTI_E_UNSTABLE instead of returning Inf/NaN seismograms."""
    shot, model = small_case(16, 20, 24, 3, 200, O.SRC_FX, False, dt=8e-3)
    with engine_for(shot, model) as e:
        e.set_probe(10, 1e6)
        with pytest.raises(qvti.QvtiError) as err:
            e.forward(shot["nt"], shot["rik"], shot_of(shot))
        assert err.value.code == qvti.E_UNSTABLE
    shot, model = small_case(16, 20, 24, 3, 200, O.SRC_FX, False)
    with engine_for(shot, model) as e:
        e.set_probe(10, 1e6)
        rec = e.forward(shot["nt"], shot["rik"], shot_of(shot))  # stable: no error
        assert np.isfinite(rec).all()


def test_reference_entry_point_matches_the_oracle():
    """qvti_forward_acoustic_3D: the reference's argument list; x-force at
    (s_ix, s_iy, s_iz + 10), receivers (r_ix[ir], r_iy[ir], r_iz) for
    ir < rnx_max * rny_max, records [nt][R] (cuviscoelas_3D.cu:771-791)."""
    shot, model = small_case(32, 24, 40, 4, 35, O.SRC_FX, False)
    nx, ny, nz, pml, nt = shot["nx"], shot["ny"], shot["nz"], shot["pml"], shot["nt"]
    rnx, rny, dr = 5, 4, 3
    rix, riy = qvti.reference_receivers(pml, rnx, rny, dr)
    ss = qvti.RefSource(8, nx // 2, ny // 2, pml + 2,
                        rix.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                        riy.ctypes.data_as(ctypes.POINTER(ctypes.c_int)), rnx * rny)
    srcs = (qvti.RefSource * 1)(ss)
    R = rnx * rny
    recs = np.zeros((3, nt, R), np.float32)
    f32 = lambda a: np.ascontiguousarray(a, np.float32)
    m = [f32(model[k]) for k in ("vp", "gp", "vs", "gs", "rho")]
    rik = f32(shot["rik"])
    with qvti.Engine(nx, ny, nz, pml, shot["dx"], shot["dy"], shot["dz"], shot["dt"], shot["f0"], shot["w0"]) as e:
        qvti.check(qvti.lib().qvti_forward_acoustic_3D(
            e._c, 0, 0, nt, nx, ny, nz, nx * ny * nz, nx - 2 * pml, ny - 2 * pml, nz - 2 * pml, pml,
            shot["dx"], shot["dy"], shot["dz"], shot["dt"], shot["f0"], shot["w0"], 4000.0,
            rik.ctypes.data, *[a.ctypes.data for a in m], ctypes.cast(srcs, ctypes.c_void_p), R, rnx, rny, dr,
            recs[0].ctypes.data, recs[1].ctypes.data, recs[2].ctypes.data))
    oshot = dict(shot, src_kind=O.SRC_FX, sx=nx // 2, sy=ny // 2, sz=8 + 10, rx=rix, ry=riy,
                 rz=np.full(R, pml + 2, np.int32))
    ref = O.run(oshot, model, "f64")["records"]
    r32 = O.run(oshot, model, "f32")["records"]
    for c in range(3):
        # vy and vz of an x-force are weak at these receivers: floor-aware bound
        floor = rel_l2(r32[c], ref[c])
        e = rel_l2(recs[c], ref[c])
        print("component %d: vs FP64 %.2e (FP32 oracle floor %.2e)" % (c, e, floor))
        assert e < max(TOL, 2 * floor), (c, e, floor)


@pytest.mark.parametrize("workload", ["vti256", "vti160"])
def test_exact_linearity_in_the_source_at_full_size(workload):
    """Scaling the wavelet by 2 scales every record bit-for-bit (every
    operation of the step is linear in the source and a factor of 2 commutes
    with FP32 rounding), at the benchmark's full 256^3 / 160x160x128 sizes; and
    batched shots equal single shots bit-for-bit there."""
    import bench

    g, m, sh = bench.make_workload(workload)
    nt = 12
    rik = qvti.ricker(nt, g["dt"], g["f0"]) * np.float32(1e3)
    shots = bench.surface_shots(g, 2, 3)
    with bench.engine(g, max_shots=2) as e:
        e.set_model_dict(m)
        a = e.forward_batch(nt, rik, shots)
        b = e.forward_batch(nt, rik * np.float32(2), shots)
    with bench.engine(g) as e:
        e.set_model_dict(m)
        single = e.forward(nt, rik, shots[1])
    assert np.abs(a[0]).max() > 0
    for x, y in zip(a, b):
        assert np.array_equal(2 * x, y)
    assert np.array_equal(single, a[1])


def test_checkpoint_restore_continues_bit_identically():
    """A run checkpointed after 17 of 40 steps and restored into a fresh
    context (records so far, snapshots, C-PML memory included) finishes with
    the same bits as the uninterrupted run, for a 2-shot batch; an image of a
    different run is refused."""
    shot, model = small_case(24, 20, 32, 4, 40, O.SRC_EXPLOSIVE, True)
    shots = [shot_of(shot), qvti.make_shot(O.SRC_FZ, (8, 9, 12), shot["rx"], shot["ry"], shot["rz"])]
    nt, k = shot["nt"], 17

    def run(split):
        with engine_for(shot, model, max_shots=2) as e:
            snaps = e.set_snapshot(4, 6, nt)
            e.begin(nt, shot["rik"], shots)
            if split:
                e.step(k)
                img = e.checkpoint()
                e.step(5)  # work after the checkpoint is discarded by the restore below
                with engine_for(shot, model, max_shots=2) as f:
                    snaps = f.set_snapshot(4, 6, nt)
                    f.begin(nt, shot["rik"], shots)
                    f.restore(img)
                    assert f.steps_done() == k
                    f.step(nt - k)
                    return f.finish(), snaps.copy(), img
            e.step(nt)
            return e.finish(), snaps.copy(), None

    (a, sa, _), (b, sb, img) = run(False), run(True)
    for x, y in zip(a, b):
        assert np.array_equal(x, y)
    assert np.array_equal(sa, sb)
    with engine_for(shot, model, max_shots=2) as e:
        e.set_snapshot(4, 6, nt)
        e.begin(nt - 1, shot["rik"], shots)
        with pytest.raises(qvti.QvtiError) as err:
            e.restore(img)
        assert err.value.code == qvti.E_ARG


def test_snapshot_window_reaches_into_the_cpml():
    """qvti_set_snapshot_window: a box that includes C-PML cells; its last
    snapshot (taken after the final step) matches the oracle's final field,
    and its interior part equals the default interior-window snapshots."""
    shot, model = small_case(24, 20, 32, 4, 31, O.SRC_EXPLOSIVE, True)
    nt, every, field = shot["nt"], 5, 3
    lo, hi = (0, 2, 1), (24, 17, 32)
    ref = O.run(shot, model, "f64", state=True)["state"][field]
    with engine_for(shot, model) as e:
        box = e.set_snapshot(field, every, nt, window=(lo, hi))
        e.begin(nt, shot["rik"], [shot_of(shot)])
        e.step(nt)
        e.finish()
        box = box.copy()
        inner = e.set_snapshot(field, every, nt)
        e.begin(nt, shot["rik"], [shot_of(shot)])
        e.step(nt)
        e.finish()
    want = ref[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]]
    assert (nt - 1) % every == 0 and box.shape == (7,) + want.shape
    assert rel_l2(box[-1], want) < TOL
    p, n = shot["pml"], (shot["nx"], shot["ny"], shot["nz"])
    part = box[:, p - lo[0]:n[0] - p - lo[0], p - lo[1]:n[1] - p - lo[1], p - lo[2]:n[2] - p - lo[2]]
    assert np.array_equal(part, inner)
    with engine_for(shot, model) as e2:
        with pytest.raises(qvti.QvtiError) as err:
            e2.set_snapshot(field, every, nt, window=((0, 0, 0), (25, 20, 32)))
        assert err.value.code == qvti.E_ARG


def test_truncated_snapshot_run_after_a_large_one_stays_in_the_host_buffer():
    """A context that ran with a large snapshot buffer, then runs with a
    smaller host capacity than its snapshots need (truncation), writes at
    most that capacity: a guard region after the host buffer stays intact,
    the floats it writes equal the start of the full run's snapshots, and a
    checkpoint of the truncated run restores into a fresh context."""
    shot, model = small_case(24, 20, 32, 4, 30, O.SRC_EXPLOSIVE, True)
    nt, every, field = shot["nt"], 5, 2
    guard = 4096
    with engine_for(shot, model) as e:
        snaps_full = e.set_snapshot(field, every, nt)
        e.begin(nt, shot["rik"], [shot_of(shot)])
        e.step(nt)
        e.finish()
        full = snaps_full.copy()
        W = full[0].size
        cap = 2 * W + W // 3  # two whole snapshots and part of a third
        buf = np.full(cap + guard, 12345.0, np.float32)
        qvti.check(qvti.lib().qvti_set_snapshot(e._c, field, every, buf.ctypes.data, cap))
        e.begin(nt, shot["rik"], [shot_of(shot)])
        e.step(nt)
        e.finish()
        assert np.all(buf[cap:] == 12345.0), "snapshot copy wrote past the host capacity"
        assert np.array_equal(buf[:2 * W], full.reshape(-1)[:2 * W])
        # checkpoint half way through a second truncated run on the reused context
        buf_a = np.full(cap + guard, 12345.0, np.float32)
        qvti.check(qvti.lib().qvti_set_snapshot(e._c, field, every, buf_a.ctypes.data, cap))
        e.begin(nt, shot["rik"], [shot_of(shot)])
        e.step(nt // 2)
        img = e.checkpoint()
    with engine_for(shot, model) as f:
        buf_b = np.full(cap + guard, 12345.0, np.float32)
        qvti.check(qvti.lib().qvti_set_snapshot(f._c, field, every, buf_b.ctypes.data, cap))
        f.begin(nt, shot["rik"], [shot_of(shot)])
        f.restore(img)
        f.step(nt - nt // 2)
        f.finish()
    assert np.all(buf_b[cap:] == 12345.0)
    assert np.array_equal(buf_b[:2 * W], full.reshape(-1)[:2 * W])


def test_step_after_a_failed_begin_reports_state_error():
    """A qvti_begin that fails after it has released the previous run's
    buffers (here: a receiver record too large for the device) leaves the
    context not steppable - qvti_step returns QVTI_E_STATE instead of
    replaying the previous run's graph - and a later valid begin works."""
    shot, model = small_case(16, 20, 24, 3, 20, O.SRC_FX, False)
    with engine_for(shot, model) as e:
        rec_ok = e.forward(shot["nt"], shot["rik"], shot_of(shot))
        e.begin(shot["nt"], shot["rik"], [shot_of(shot)])
        e.step(5)
        e.sync()
        nrec = 1_000_000  # 3 * nrec * nt floats: about 1.2 TB of records
        big = qvti.make_shot(O.SRC_FX, (shot["sx"], shot["sy"], shot["sz"]), np.zeros(nrec, np.int32),
                             np.zeros(nrec, np.int32), np.zeros(nrec, np.int32))
        nt_big = 100_000
        rik_big = np.zeros(nt_big, np.float32)
        with pytest.raises(qvti.QvtiError) as err:
            e.begin(nt_big, rik_big, [big])
        assert err.value.code == qvti.E_NOMEM
        with pytest.raises(qvti.QvtiError) as err:
            e.step(1)
        assert err.value.code == qvti.E_STATE
        rec = e.forward(shot["nt"], shot["rik"], shot_of(shot))
        assert np.array_equal(rec, rec_ok)
