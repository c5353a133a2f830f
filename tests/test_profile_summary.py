"""CPU test of bench.py's rocprofv3 summarizer: dispatches are assigned to the
step's kernel classes by launch order, and the tile tuner's single-class
launches around the steps are left out."""
import os
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def _name(base):
    return "void qvti::%s<qvti::L2<16, 16, 2>, 0>(qvti::Args)" % base


def test_step_classes_follow_launch_order():
    sig = [k for k, _ in bench.STEP_SIGNATURE]
    disp = []
    # tuner: each class alone, warm-up + 3 repetitions
    for base in ("k2_deriv_strided", "k2_cfft_strided", "k2_xpass"):
        disp += [(len(disp), _name(base), -1.0)] * 4
    for step in range(2):
        for j, base in enumerate(sig):
            disp.append((len(disp), _name(base), 100.0 * step + j))
    disp.append((len(disp), _name("k2_deriv_strided"), -1.0))  # a lone launch after the steps
    out = bench._step_classes(disp)
    for j, (_, cls) in enumerate(bench.STEP_SIGNATURE):
        assert out[cls] == [float(j), 100.0 + j], cls


def test_step_classes_ignore_incomplete_steps():
    sig = [k for k, _ in bench.STEP_SIGNATURE]
    disp = [(i, _name(b), 1.0) for i, b in enumerate(sig[:-1])]  # no record launch
    assert all(v == [] for v in bench._step_classes(disp).values())
