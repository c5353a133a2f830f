"""GPU parity at the benchmark configurations' full sizes against oracle
fixtures computed in the build container (tests/golden/make_fullsize_golden.py,
cases in tests/fullsize_cases.py).

* 256^3 and 400x400x200 smooth random VTI single shots (the headline and the
  non-power-of-two configuration), 200 and 160 steps of 0.5 ms: the
  wavefront reaches receivers 10-40 cells from the source and enters the
  C-PML; no receiver is the source cell.
* The 32-shot batch on one 160x160x128 model: every shot of the batched run
  is bit-identical to its single-shot run, and shot 0 matches its fixture.

Bound: relative L2 <= 1e-5 over the whole gather (3 components x nt x R)
against both the FP32 oracle (the stand-in for the reference's own FP32 run)
and the FP64 oracle.
"""
import os

import numpy as np
import pytest

import fullsize_cases as FC
from cases import rel_l2

pytestmark = pytest.mark.gpu
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
TOL = 1e-5


def _report(label, rec, gold, near_mask=None):
    e32 = rel_l2(rec, gold["rec_f32"])
    e64 = rel_l2(rec, gold["rec_f64"])
    floor = rel_l2(gold["rec_f32"], gold["rec_f64"])
    comps = ", ".join("%s %.2e" % (n, rel_l2(rec[c], gold["rec_f64"][c])) for c, n in enumerate(("vx", "vy", "vz")))
    msg = "%s: whole gather vs FP32 oracle %.3e, vs FP64 oracle %.3e (FP32 oracle vs FP64 %.3e); vs FP64 per component: %s" % (
        label, e32, e64, floor, comps)
    far = None
    if near_mask is not None:
        share = np.linalg.norm(gold["rec_f64"][:, :, ~near_mask]) ** 2 / np.linalg.norm(gold["rec_f64"]) ** 2
        f = ~near_mask
        far = (rel_l2(rec[:, :, f], gold["rec_f32"][:, :, f]), rel_l2(rec[:, :, f], gold["rec_f64"][:, :, f]),
               rel_l2(gold["rec_f32"][:, :, f], gold["rec_f64"][:, :, f]))
        msg += ("; share of the gather's energy at receivers >= 10 cells from the source: %.1f %%; those receivers "
                "alone: vs FP32 %.3e, vs FP64 %.3e (FP32 oracle vs FP64 %.3e)" % ((100 * share,) + far))
    print(msg, flush=True)
    return e32, e64, far


@pytest.mark.parametrize("name", sorted(FC.FULLSIZE))
def test_full_size_single_shot_matches_the_oracle_fixture(name):
    gold = np.load(os.path.join(GOLDEN, "fullsize_%s.npz" % name))
    g, m, shot = FC.fullsize_case(name)
    assert FC.model_checksum(m) == gold["model_sum"]
    assert np.array_equal(shot["rik"], gold["rik"]) and np.array_equal(shot["rx"], gold["rx"])
    print("case %s: %dx%dx%d, %d steps, %d receivers" % (name, g["nx"], g["ny"], g["nz"], shot["nt"], len(shot["rx"])),
          flush=True)
    with FC.engine_for(g) as e:
        e.set_model_dict(m)
        rec = e.forward(shot["nt"], shot["rik"], FC.make_shot(shot))
    d = np.sqrt((shot["rx"] - shot["src"][0]) ** 2.0 + (shot["ry"] - shot["src"][1]) ** 2.0 +
                (shot["rz"] - shot["src"][2]) ** 2.0)
    e32, e64, far = _report(name, rec, gold, near_mask=d < 10)
    assert np.all(np.isfinite(rec))
    assert e32 < TOL and e64 < TOL
    assert far[0] < TOL and far[1] < TOL  # the receivers 10-40 cells out and in the C-PML on their own


def test_32_shot_batch_equals_single_shots_and_matches_the_fixture():
    gold = np.load(os.path.join(GOLDEN, "batch_vti160.npz"))
    g, m, shots = FC.batch_case()
    assert FC.model_checksum(m) == gold["model_sum"]
    assert tuple(gold["src"]) == shots[0]["src"] and np.array_equal(gold["rx"], shots[0]["rx"])
    nt, rik = FC.BATCH_NT, FC.wavelet(FC.BATCH_NT)
    print("case batch: %d shots of %dx%dx%d, %d steps" % (len(shots), g["nx"], g["ny"], g["nz"], nt), flush=True)
    with FC.engine_for(g, max_shots=len(shots)) as e:
        e.set_model_dict(m)
        batch = e.forward_batch(nt, rik, [FC.make_shot(s) for s in shots])
    with FC.engine_for(g) as e:
        e.set_model_dict(m)
        for k, s in enumerate(shots):
            single = e.forward(nt, rik, FC.make_shot(s))
            assert np.array_equal(batch[k], single), "shot %d differs between batched and single runs" % k
    e32, e64, _ = _report("32-shot batch, shot 0", batch[0], gold)
    assert e32 < TOL and e64 < TOL
