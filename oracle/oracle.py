"""ctypes wrapper of the CPU oracle (oracle/qvti_oracle.c).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and
bench.py's cpu_baseline leg, to check the HIP engine's results and to time
the CPU baseline.  The product path never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libqvti_oracle.so")

SRC_FX, SRC_FY, SRC_FZ, SRC_EXPLOSIVE = 0, 1, 2, 3
FIELDS = ["vx", "vy", "vz", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz"]


class QoConfig(ctypes.Structure):
    _fields_ = [
        ("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int),
        ("pml", ctypes.c_int), ("nt", ctypes.c_int),
        ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float),
        ("dt", ctypes.c_float), ("f0", ctypes.c_float), ("w0", ctypes.c_float),
        ("src_kind", ctypes.c_int), ("sx", ctypes.c_int), ("sy", ctypes.c_int), ("sz", ctypes.c_int),
        ("nrec", ctypes.c_int),
        ("rx", ctypes.POINTER(ctypes.c_int)), ("ry", ctypes.POINTER(ctypes.c_int)),
        ("rz", ctypes.POINTER(ctypes.c_int)),
        ("vti", ctypes.c_int), ("snap_every", ctypes.c_int), ("snap_field", ctypes.c_int),
        ("nthreads", ctypes.c_int),
        ("ks_vref", ctypes.c_float), ("ks_fcut", ctypes.c_float), ("ks_vmax", ctypes.c_float),
        ("ks_taper", ctypes.c_float), ("slip", ctypes.c_int),
        ("ninj", ctypes.c_int), ("inj_x", ctypes.POINTER(ctypes.c_int)), ("inj_y", ctypes.POINTER(ctypes.c_int)),
        ("inj_z", ctypes.POINTER(ctypes.c_int)), ("inj_vx", ctypes.c_void_p), ("inj_vy", ctypes.c_void_p),
        ("inj_vz", ctypes.c_void_p), ("hist_field", ctypes.c_int), ("hist", ctypes.c_void_p),
        ("media", ctypes.c_int),
    ]


PRESSURE = 9  # history field (pxx + pyy + pzz) / 3


# planted mistakes (slip tests only): qvti_oracle.h QO_SLIP_*
SLIP_ETA_SWAP, SLIP_NO_TAU, SLIP_SHEAR_DPLUS, SLIP_PML_NODE_V, SLIP_PML_NODE_S, SLIP_ETA_SIGN = 1, 2, 4, 8, 16, 32


_lib = None


def build():
    subprocess.check_call(["make", "-s", "-C", HERE])


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        _lib = ctypes.CDLL(LIB_PATH)
        fp = np.ctypeslib.ndpointer
        for sfx, dt in (("f32", np.float32), ("f64", np.float64)):
            f = getattr(_lib, "qo_fft1d_" + sfx)
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, fp(dt, flags="C")]
            f = getattr(_lib, "qo_fft3d_" + sfx)
            f.argtypes = [ctypes.c_int] * 4 + [fp(dt, flags="C")]
            f = getattr(_lib, "qo_axis_tables_" + sfx)
            f.argtypes = [ctypes.c_int, ctypes.c_float] + [fp(dt, flags="C")] * 3
            f = getattr(_lib, "qo_pml_profile_" + sfx)
            f.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_float, ctypes.c_float] + [fp(dt, flags="C")] * 2
            f = getattr(_lib, "qo_run_" + sfx)
            f.argtypes = [ctypes.POINTER(QoConfig)] + [ctypes.c_void_p] * 13
            f = getattr(_lib, "qo_kspace_factor_" + sfx)
            f.argtypes = [ctypes.c_int, fp(dt, flags="C")] + [ctypes.c_float] * 5 + [fp(dt, flags="C")]
        _lib.qo_q_coeffs.argtypes = [ctypes.c_double] * 4 + [fp(np.float64, flags="C")]
    return _lib


def fft1d(x, sign=-1, prec="f64"):
    """Unnormalised FFT along the last axis (sign -1 forward)."""
    dt = np.float64 if prec == "f64" else np.float32
    x = np.ascontiguousarray(x, dtype=np.complex128 if prec == "f64" else np.complex64)
    n = x.shape[-1]
    buf = x.view(dt).reshape(-1).copy()
    rc = getattr(lib(), "qo_fft1d_" + prec)(n, sign, x.size // n, buf)
    assert rc == 0
    return buf.view(x.dtype).reshape(x.shape)


def fft3d(x, sign=-1, prec="f64"):
    dt = np.float64 if prec == "f64" else np.float32
    x = np.ascontiguousarray(x, dtype=np.complex128 if prec == "f64" else np.complex64)
    buf = x.view(dt).reshape(-1).copy()
    rc = getattr(lib(), "qo_fft3d_" + prec)(x.shape[0], x.shape[1], x.shape[2], sign, buf)
    assert rc == 0
    return buf.view(x.dtype).reshape(x.shape)


def axis_tables(n, d, prec="f64"):
    dt = np.float64 if prec == "f64" else np.float32
    k = np.zeros(n, dt)
    dp = np.zeros(2 * n, dt)
    dm = np.zeros(2 * n, dt)
    assert getattr(lib(), "qo_axis_tables_" + prec)(n, d, k, dp, dm) == 0
    return k, dp[0::2] + 1j * dp[1::2], dm[0::2] + 1j * dm[1::2]


def pml_profile(n, pml, d, f0, prec="f64"):
    dt = np.float64 if prec == "f64" else np.float32
    a = np.zeros(n, dt)
    b = np.zeros(n, dt)
    assert getattr(lib(), "qo_pml_profile_" + prec)(n, pml, d, f0, a, b) == 0
    return a, b


def q_coeffs(rho, v, gamma, w0):
    out = np.zeros(5, np.float64)
    lib().qo_q_coeffs(rho, v, gamma, w0, out)
    return out


def kspace_factor(kabs, dt, vref=0.0, fcut=0.0, vmax=0.0, taper=0.0, prec="f64"):
    """The oracle's k-space factor F(|k|) of the stress operators."""
    dt_ = np.float64 if prec == "f64" else np.float32
    k = np.ascontiguousarray(kabs, dt_).ravel()
    out = np.zeros_like(k)
    rc = getattr(lib(), "qo_kspace_factor_" + prec)(len(k), k, dt, vref, fcut, vmax, taper, out)
    if rc != 0:
        raise ValueError("oracle rejected the k-space parameters")
    return out.reshape(np.shape(kabs))


def _ptr(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def run(shot, model, prec="f32", nthreads=0, state=False, snap_every=0, snap_field=0, kspace=None, slip=0,
        inject=None, hist_field=None, media=0):
    """Run the oracle on one shot.

    shot: dict(nx, ny, nz, pml, nt, dx, dy, dz, dt, f0, w0, src_kind, sx, sy, sz,
               rx, ry, rz (int arrays), rik (float32 [nt]))
    model: dict(vp, gp, vs, gs, rho[, eps, delta]) float32 arrays of shape (nx, ny, nz)
    kspace: optional dict(vref, fcut, vmax, taper) of the opt-in k-space
    correction / wavenumber filter (absent: the reference scheme).
    slip: SLIP_* bits of a planted mistake (slip tests only; 0: none).
    inject: optional dict(x, y, z (int arrays [R]), vx, vy, vz (float32
    [nt][R] or None)) of extra point sources added to v after the shot's
    source each step; hist_field: a field id or PRESSURE whose whole grid is
    returned at every step as out["hist"] [nt][nx][ny][nz]; media: 0
    constant-Q, 1 dispersion only, 2 lossless, 3 Q-compensating.
    Returns dict with records [3][nt][nrec] and optionally state [9][nx][ny][nz]
    and snapshots.
    """
    dt = np.float64 if prec == "f64" else np.float32
    nx, ny, nz, nt = shot["nx"], shot["ny"], shot["nz"], shot["nt"]
    rx = np.ascontiguousarray(shot["rx"], np.int32)
    ry = np.ascontiguousarray(shot["ry"], np.int32)
    rz = np.ascontiguousarray(shot["rz"], np.int32)
    nrec = len(rx)
    vti = int("eps" in model and model["eps"] is not None)
    cfg = QoConfig(nx, ny, nz, shot["pml"], nt, shot["dx"], shot["dy"], shot["dz"], shot["dt"],
                   shot["f0"], shot["w0"], shot["src_kind"], shot["sx"], shot["sy"], shot["sz"], nrec,
                   rx.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                   ry.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                   rz.ctypes.data_as(ctypes.POINTER(ctypes.c_int)),
                   vti, snap_every, snap_field, nthreads)
    cfg.slip = int(slip)
    cfg.media = int(media)
    keep = []
    if inject is not None:
        ix = [np.ascontiguousarray(inject[k], np.int32) for k in ("x", "y", "z")]
        keep += ix
        cfg.ninj = len(ix[0])
        cfg.inj_x, cfg.inj_y, cfg.inj_z = [a.ctypes.data_as(ctypes.POINTER(ctypes.c_int)) for a in ix]
        for name in ("vx", "vy", "vz"):
            tr = inject.get(name)
            if tr is not None:
                tr = np.ascontiguousarray(tr, np.float32)
                assert tr.shape == (nt, cfg.ninj)
                keep.append(tr)
                setattr(cfg, "inj_" + name, tr.ctypes.data)
    hist = None
    cfg.hist_field = -1
    if hist_field is not None:
        hist = np.zeros((nt, nx, ny, nz), dt)
        cfg.hist_field = int(hist_field)
        cfg.hist = hist.ctypes.data
    if kspace:
        cfg.ks_vref = kspace.get("vref", 0.0)
        cfg.ks_fcut = kspace.get("fcut", 0.0)
        cfg.ks_vmax = kspace.get("vmax", 0.0)
        cfg.ks_taper = kspace.get("taper", 0.0)
    f32 = lambda a: None if a is None else np.ascontiguousarray(a, np.float32)
    rik = f32(shot["rik"])
    m = {k: f32(model.get(k)) for k in ("vp", "gp", "vs", "gs", "rho", "eps", "delta")}
    recs = np.zeros((3, nt, nrec), dt)
    st = np.zeros((9, nx, ny, nz), dt) if state else None
    snaps = None
    if snap_every:
        ns = (nt + snap_every - 1) // snap_every
        p = shot["pml"]
        snaps = np.zeros((ns, nx - 2 * p, ny - 2 * p, nz - 2 * p), dt)
    rc = getattr(lib(), "qo_run_" + prec)(
        ctypes.byref(cfg), _ptr(rik), _ptr(m["vp"]), _ptr(m["gp"]), _ptr(m["vs"]), _ptr(m["gs"]),
        _ptr(m["rho"]), _ptr(m["eps"]), _ptr(m["delta"]),
        _ptr(recs[0]), _ptr(recs[1]), _ptr(recs[2]), _ptr(snaps), _ptr(st))
    if rc != 0:
        raise ValueError("oracle rejected the configuration (rc=%d)" % rc)
    out = {"records": recs}
    if state:
        out["state"] = st
    if snap_every:
        out["snaps"] = snaps
    if hist is not None:
        out["hist"] = hist
    return out
