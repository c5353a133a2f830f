"""CPU tests of the C ABI library (no GPU): it loads, exports every entry point
include/qvti.h declares, fails loudly without a device, and its host helpers
(viscoelas_3D.cpp's wavelet, gamma and receiver geometry; the seeded model
generators) match their closed forms."""
import ctypes
import os
import re

import numpy as np
import pytest

import qvti

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
PI_T = 3.1415926


def header_functions():
    text = open(os.path.join(ROOT, "include", "qvti.h")).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(qvti_[A-Za-z0-9_]+)\s*\(", text)))


def test_library_exports_every_declared_entry_point():
    L = qvti.lib()
    names = header_functions()
    assert len(names) >= 30
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    assert set(names) == set(qvti.EXPORTS)


def test_no_device_fails_loudly():
    if qvti.device_count() > 0:
        pytest.skip("a device is visible")
    c = ctypes.c_void_p()
    g = qvti.Grid(16, 16, 16, 3, 10, 10, 10, 1e-3, 25, 157, 0)
    rc = qvti.lib().qvti_create(ctypes.byref(c), ctypes.byref(g), 1, 0)
    assert rc == qvti.E_DEVICE
    assert b"device" in qvti.lib().qvti_last_error().lower()


def _smooth_even(n):
    if n % 2:
        return False
    for p in (2, 3, 5):
        while n % p == 0:
            n //= p
    return n == 1


def test_unsupported_lengths_are_rejected_before_the_device():
    """Axis lengths outside the even 2/3/5-smooth numbers of [16, 512] (a
    factor 7 or 97, odd, too short, too long) are refused before any device
    work, and the message lists exactly the even 2/3/5-smooth lengths of
    [16, 512] as the supported ones (cuFFT plans any smooth size for the
    reference, cuviscoelas_3D.cu:1703-1704)."""
    c = ctypes.c_void_p()
    for dims in ((14, 16, 16), (16, 16, 49), (535, 485, 48), (16, 10, 16), (16, 16, 1024), (194, 16, 16)):
        g = qvti.Grid(*dims, 3, 10, 10, 10, 1e-3, 25, 157, 0)
        rc = qvti.lib().qvti_create(ctypes.byref(c), ctypes.byref(g), 1, 0)
        assert rc == qvti.E_UNSUPPORTED, dims
        msg = qvti.lib().qvti_last_error().decode()
        listed = [int(t) for t in msg.split("supported axis lengths are")[1].split()]
        assert listed == [n for n in range(16, 513) if _smooth_even(n)]


def test_ricker_matches_reference_formula():
    """viscoelas_3D.cpp:498-502: tmp = (pi f0 (it dt - 1/f0))^2 held in FP32."""
    nt, dt, f0 = 400, np.float32(1e-3), np.float32(25.0)
    r = qvti.ricker(nt, float(dt), float(f0))
    it = np.arange(nt)
    x = (it * dt).astype(np.float32) - np.float32(1) / f0
    tmp = ((PI_T * float(f0)) * x.astype(np.float64)) ** 2
    tmp = tmp.astype(np.float32)
    ref = ((1.0 - 2.0 * tmp.astype(np.float64)) * np.exp(-tmp).astype(np.float32)).astype(np.float32)
    assert np.allclose(r, ref, rtol=3e-7, atol=1e-12)  # exp implementations differ by <= 1 ulp
    assert abs(r.max() - 1.0) < 1e-6 and np.argmax(r) == 40


def test_gamma_and_receivers():
    q = np.array([20, 50, 100, 200, 1e6], np.float32)
    assert np.allclose(qvti.gamma_from_q(q), (np.arctan(1.0 / q.astype(np.float64)) / PI_T).astype(np.float32), rtol=0)
    rx, ry = qvti.reference_receivers(15, 4, 3, 2)
    ip = np.arange(12)
    assert np.array_equal(rx, 15 + ip // 3 * 2) and np.array_equal(ry, 15 + ip % 3 * 2)


def test_model_generators():
    shape = (20, 16, 24)
    h = qvti.make_model(0, 1, shape, [3000, 1730, 2200, 100, 50])
    assert np.all(h["vp"] == 3000) and np.all(h["qs"] == 50) and np.all(h["rho"] == 2200)
    lay = qvti.make_model(1, 7, shape, [1800, 4500, 0.25, 30, 200], nlay=6, vti=True)
    assert len(np.unique(lay["vp"])) == 6
    assert np.all(lay["vp"] == lay["vp"][:1, :1, :])          # flat in x and y
    assert lay["vp"].min() >= 1800 and lay["vp"].max() <= 4500
    assert np.allclose(lay["vs"], lay["vp"] / 1.73, rtol=1e-6)
    assert np.allclose(lay["rho"], 310 * lay["vp"].astype(np.float64) ** 0.25, rtol=1e-6)
    assert np.all(lay["delta"] <= lay["eps"]) and lay["eps"].max() <= 0.25
    assert np.allclose(lay["qs"], lay["qp"] / 2)
    sm = qvti.make_model(2, 3, shape, [2000, 5000, 0.25, 20, 150, 3, 0.05], vti=True)
    assert abs(sm["vp"].min() - 2000) < 1e-3 and abs(sm["vp"].max() - 5000) < 1e-2
    assert sm["qp"].min() >= 20 - 1e-4 and sm["qp"].max() <= 150 + 1e-3
    assert np.allclose(sm["delta"], sm["eps"] / 2, rtol=1e-6)
    # eps decreases as vp increases (co-varying fields)
    assert sm["eps"].ravel()[np.argmax(sm["vp"])] < sm["eps"].ravel()[np.argmin(sm["vp"])]
    again = qvti.make_model(2, 3, shape, [2000, 5000, 0.25, 20, 150, 3, 0.05], vti=True)
    assert np.array_equal(again["vp"], sm["vp"])


def test_estimate_bytes_scales_with_shots():
    g = qvti.Grid(256, 256, 256, 20, 12.5, 12.5, 12.5, 1e-3, 20, 125.7, 1)
    one = qvti.lib().qvti_estimate_bytes(ctypes.byref(g), 1)
    four = qvti.lib().qvti_estimate_bytes(ctypes.byref(g), 4)
    assert 3e9 < one < 12e9 and four > 2.5 * one


def _reference_padding(interior, pml):
    """The reference's padding loops, restated literally (viscoelas_3D.cpp:576-655):
    z for interior x, y; then y for interior x, all z; then x for all y, z."""
    wx, wy, wz = interior.shape
    ntx, nty, ntz = wx + 2 * pml, wy + 2 * pml, wz + 2 * pml
    v = np.zeros((ntx, nty, ntz), np.float32)
    v[pml:ntx - pml, pml:nty - pml, pml:ntz - pml] = interior
    for ix in range(pml, ntx - pml):
        for iy in range(pml, nty - pml):
            for iz in range(pml):
                v[ix, iy, iz] = v[ix, iy, pml]
            for iz in range(ntz - pml, ntz):
                v[ix, iy, iz] = v[ix, iy, ntz - pml - 1]
    for iz in range(ntz):
        for ix in range(pml, ntx - pml):
            for iy in range(pml):
                v[ix, iy, iz] = v[ix, pml, iz]
            for iy in range(nty - pml, nty):
                v[ix, iy, iz] = v[ix, nty - pml - 1, iz]
    for iz in range(ntz):
        for iy in range(nty):
            for ix in range(pml):
                v[ix, iy, iz] = v[pml, iy, iz]
            for ix in range(ntx - pml, ntx):
                v[ix, iy, iz] = v[ntx - pml - 1, iy, iz]
    return v


def test_model_padding_matches_the_reference_loops():
    rng = np.random.default_rng(4)
    for shape, pml in (((5, 7, 6), 3), ((4, 4, 9), 2), ((1, 2, 3), 4)):
        interior = rng.random(shape).astype(np.float32)
        assert np.array_equal(qvti.pad_model(interior, pml), _reference_padding(interior, pml)), (shape, pml)


def test_raw_model_and_record_files_round_trip(tmp_path):
    rng = np.random.default_rng(5)
    interior = rng.random((6, 5, 7)).astype(np.float32)
    path = str(tmp_path / "vp.bin")
    interior.tofile(path)  # the reference's input format: raw FP32, ix slowest
    total = qvti.load_model_raw(path, (12, 11, 13), 3)
    assert np.array_equal(total, _reference_padding(interior, 3))
    with pytest.raises(qvti.QvtiError):
        qvti.load_model_raw(path, (14, 11, 13), 3)  # file too short for that interior
    rec = rng.random((40, 9)).astype(np.float32)  # [nt][R]
    out = str(tmp_path / "rec.bin")
    qvti.write_raw(out, rec)
    assert np.array_equal(np.fromfile(out, np.float32).reshape(40, 9), rec)


def test_split_wisdom_file_round_trip(tmp_path):
    """qvti_load_wisdom merges "nx ny nz vti split_x split_y split_z" lines
    into the process table that qvti_create consults, qvti_save_wisdom
    writes it back; a split variant a length does not have, or a malformed
    file, is refused (no device needed)."""
    good = tmp_path / "w.txt"
    good.write_text("qvti-wisdom 1\n256 256 256 1 1 0 1\n400 400 200 1 0 1 0\n")
    qvti.load_wisdom(str(good))
    out = tmp_path / "out.txt"
    qvti.save_wisdom(str(out))
    lines = out.read_text().split("\n")
    assert lines[0] == "qvti-wisdom 1"
    assert "256 256 256 1 1 0 1" in lines and "400 400 200 1 0 1 0" in lines
    bad = tmp_path / "bad.txt"
    bad.write_text("qvti-wisdom 1\n90 90 90 0 1 0 0\n")  # 90 has no alternative split
    with pytest.raises(qvti.QvtiError) as err:
        qvti.load_wisdom(str(bad))
    assert err.value.code == qvti.E_ARG
    bad.write_text("not wisdom\n")
    with pytest.raises(qvti.QvtiError):
        qvti.load_wisdom(str(bad))
