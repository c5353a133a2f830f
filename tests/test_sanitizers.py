"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

The CPU oracle (oracle/qvti_oracle.c) and the engine library's host helpers
(q-vti_amd/csrc/qvti_host.cpp, qvti_error.cpp: wavelet, gamma, receiver
geometry, seeded models, C-PML padding, raw file I/O, error reporting) are
rebuilt with gcc/g++ and -fsanitize=address,undefined into a temporary
folder and driven from a child Python process that preloads the ASan
runtime: every call of the oracle's FP32/FP64 runs (snapshots, state,
k-space mode, VTI, all source kinds, transforms, argument errors) and of the
helpers (including their error paths) must finish without a report.  The
GPU kernels are not covered: device sanitizers are not available for them.
"""
import os
import shutil
import subprocess
import sys
import textwrap

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]

CHILD = r"""
import ctypes, os, sys
import numpy as np
sys.path.insert(0, os.path.join(ROOT, "oracle")); sys.path.insert(0, os.path.join(ROOT, "tests"))
import oracle as O
O.LIB_PATH = ORACLE; O._lib = None
from cases import small_case
for kind, vti in ((O.SRC_FX, False), (O.SRC_EXPLOSIVE, True), (O.SRC_FZ, True), (O.SRC_FY, False)):
    shot, model = small_case(12, 10, 16, 3, 6, kind, vti)
    for prec in ("f32", "f64"):
        O.run(shot, model, prec, state=True, snap_every=2, snap_field=4, nthreads=2)
        O.run(shot, model, prec, kspace=dict(vref=3500.0, fcut=60.0, vmax=3000.0, taper=0.3))
try:
    O.run(dict(shot, sx=999), model, "f32")
    raise SystemExit("bad source accepted")
except ValueError:
    pass
x = (np.random.default_rng(0).standard_normal((6, 10, 12)) + 0j)
for prec in ("f32", "f64"):
    O.fft3d(x, -1, prec); O.fft1d(x, 1, prec)
    O.axis_tables(20, 12.5, prec); O.pml_profile(24, 4, 10.0, 25.0, prec)
    O.kspace_factor(np.linspace(0, 0.5, 50), 1e-3, 3000.0, 50.0, 3000.0, 0.2, prec)
O.q_coeffs(2000.0, 3000.0, 0.01, 150.0)

H = ctypes.CDLL(HOST)
fp = lambda a: a.ctypes.data_as(ctypes.c_void_p)
rik = np.zeros(100, np.float32)
assert H.qvti_ricker(100, ctypes.c_float(1e-3), ctypes.c_float(20.0), fp(rik)) == 0
q = np.linspace(10, 200, 64).astype(np.float32); g = np.zeros_like(q)
assert H.qvti_gamma_from_q(fp(q), fp(g), ctypes.c_longlong(64)) == 0
rx = np.zeros(12, np.int32); ry = np.zeros(12, np.int32)
assert H.qvti_reference_receivers(5, 3, 4, 2, fp(rx), fp(ry)) == 0
shape = (20, 16, 24)
N = 20 * 16 * 24
for kind, nlay, params in ((0, 1, [3000, 1730, 2200, 100, 50]), (1, 4, [1800, 4500, 0.25, 30, 200]),
                           (2, 1, [2000, 5000, 0.25, 20, 150, 4.0, 0.05, 0])):
    p = np.zeros(8, np.float64); p[:len(params)] = params
    out = [np.zeros(N, np.float32) for _ in range(7)]
    rc = H.qvti_make_model(kind, ctypes.c_uint64(7), 20, 16, 24, nlay, fp(p), *[fp(o) for o in out])
    assert rc == 0, kind
pml = 4
interior = np.random.default_rng(1).random((20 - 2 * pml) * (16 - 2 * pml) * (24 - 2 * pml)).astype(np.float32)
total = np.zeros(N, np.float32)
assert H.qvti_pad_model(fp(interior), 20, 16, 24, pml, fp(total)) == 0
path = os.path.join(TMP, "m.raw").encode()
assert H.qvti_write_raw(path, fp(interior), ctypes.c_longlong(interior.size)) == 0
total2 = np.zeros(N, np.float32)
assert H.qvti_load_model_raw(path, 20, 16, 24, pml, fp(total2)) == 0
assert np.array_equal(total, total2)
# error paths: bad sizes, missing file, NULL pointers
assert H.qvti_pad_model(fp(interior), 20, 16, 24, 11, fp(total)) != 0
assert H.qvti_load_model_raw(os.path.join(TMP, "missing.raw").encode(), 20, 16, 24, pml, fp(total2)) != 0
assert H.qvti_write_raw(path, None, ctypes.c_longlong(5)) != 0
H.qvti_last_error.restype = ctypes.c_char_p
assert H.qvti_last_error()
print("sanitized host code: ok")
"""


def _runtime(name):
    p = subprocess.run(["gcc", "-print-file-name=" + name], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.skipif(shutil.which("gcc") is None or shutil.which("g++") is None, reason="needs gcc and g++")
def test_oracle_and_host_helpers_under_asan_ubsan(tmp_path):
    asan = _runtime("libasan.so")
    if asan is None:
        pytest.skip("gcc has no ASan runtime")
    oracle_so = str(tmp_path / "liboracle_san.so")
    host_so = str(tmp_path / "libhost_san.so")
    subprocess.run(["gcc"] + SAN + ["-fopenmp", "-fPIC", "-std=c11", "-shared", "-o", oracle_so,
                                    os.path.join(ROOT, "oracle", "qvti_oracle.c"), "-lm"], check=True)
    csrc = os.path.join(ROOT, "q-vti_amd", "csrc")
    subprocess.run(["g++"] + SAN + ["-fPIC", "-std=c++17", "-shared", "-I", os.path.join(ROOT, "include"), "-I", csrc,
                                    "-o", host_so, os.path.join(csrc, "qvti_host.cpp"),
                                    os.path.join(csrc, "qvti_error.cpp"), "-lpthread"], check=True)
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    code = "ROOT, ORACLE, HOST, TMP = %r, %r, %r, %r\n" % (ROOT, oracle_so, host_so, str(tmp_path)) + textwrap.dedent(CHILD)
    r = subprocess.run([sys.executable, "-c", code], env=env, capture_output=True, text=True, timeout=600)
    report = r.stdout[-3000:] + r.stderr[-6000:]
    assert r.returncode == 0, report
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, report
    assert "sanitized host code: ok" in r.stdout
