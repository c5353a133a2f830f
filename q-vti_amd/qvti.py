"""Python binding of the engine's C ABI (include/qvti.h) over ctypes.

Mirrors the reference's host interface (viscoelas_3D.cpp: model arrays over
the total grid, Source geometry, three records [nt][R]) for tests and the
benchmark.  The library is the in-tree q-vti_amd/libqvti.so; importing this
module fails loudly when it is missing (there is no CPU fallback).
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("QVTI_LIB") or os.path.join(HERE, "libqvti.so")

OK = 0
E_ARG, E_DEVICE, E_NOMEM, E_UNSUPPORTED, E_STATE, E_UNSTABLE, E_BUDGET, E_CANCELLED, E_PENDING = range(-1, -10, -1)
SRC_FX, SRC_FY, SRC_FZ, SRC_EXPLOSIVE = 0, 1, 2, 3
FIELDS = ["vx", "vy", "vz", "pxx", "pyy", "pzz", "pxy", "pxz", "pyz"]

EXPORTS = [
    "qvti_device_count", "qvti_create", "qvti_set_model", "qvti_forward", "qvti_forward_batch",
    "qvti_forward_acoustic_3D", "qvti_set_snapshot", "qvti_destroy", "qvti_last_error", "qvti_begin",
    "qvti_step", "qvti_sync", "qvti_finish", "qvti_get_state", "qvti_steps_done",
    "qvti_enable_kernel_timing", "qvti_kernel_times", "qvti_kernel_bytes", "qvti_set_graphs",
    "qvti_set_probe", "qvti_device_bytes", "qvti_estimate_bytes", "qvti_fft3d", "qvti_deriv1d",
    "qvti_ricker", "qvti_gamma_from_q", "qvti_reference_receivers", "qvti_make_model",
    "qvti_server_create", "qvti_server_submit", "qvti_server_poll", "qvti_server_wait",
    "qvti_server_cancel", "qvti_server_clock", "qvti_server_destroy", "qvti_set_autotune", "qvti_tile_choice",
    "qvti_pad_model", "qvti_load_model_raw", "qvti_write_raw", "qvti_set_kspace",
    "qvti_checkpoint_bytes", "qvti_checkpoint", "qvti_restore", "qvti_set_snapshot_window",
    "qvti_fft3d_noalloc", "qvti_batch_width", "qvti_forward_shots", "qvti_set_media", "qvti_rtm",
    "qvti_set_split", "qvti_get_split", "qvti_tune_split", "qvti_save_wisdom", "qvti_load_wisdom",
]


class QvtiError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__("qvti error %d: %s" % (code, msg))
        self.code = code


class Grid(ctypes.Structure):
    _fields_ = [("nx", ctypes.c_int), ("ny", ctypes.c_int), ("nz", ctypes.c_int), ("pml", ctypes.c_int),
                ("dx", ctypes.c_float), ("dy", ctypes.c_float), ("dz", ctypes.c_float),
                ("dt", ctypes.c_float), ("f0", ctypes.c_float), ("w0", ctypes.c_float), ("vti", ctypes.c_int)]


class Shot(ctypes.Structure):
    _fields_ = [("kind", ctypes.c_int), ("sx", ctypes.c_int), ("sy", ctypes.c_int), ("sz", ctypes.c_int),
                ("nrec", ctypes.c_int), ("rx", ctypes.POINTER(ctypes.c_int)),
                ("ry", ctypes.POINTER(ctypes.c_int)), ("rz", ctypes.POINTER(ctypes.c_int))]


class RefSource(ctypes.Structure):
    """Layout-identical to the reference's struct Source (Myfunctions.h:6-10)."""
    _fields_ = [("s_iz", ctypes.c_int), ("s_ix", ctypes.c_int), ("s_iy", ctypes.c_int), ("r_iz", ctypes.c_int),
                ("r_ix", ctypes.POINTER(ctypes.c_int)), ("r_iy", ctypes.POINTER(ctypes.c_int)),
                ("r_n", ctypes.c_int)]


class Request(ctypes.Structure):
    _fields_ = [("grid", Grid), ("nt", ctypes.c_int)] + [
        (k, ctypes.c_void_p) for k in ("velp", "gama_p", "vels", "gama_s", "rho", "eps", "delta", "rik")
    ] + [("shot", Shot)]


class ResultInfo(ctypes.Structure):
    _fields_ = [("status", ctypes.c_int), ("t_submit_s", ctypes.c_double), ("t_admit_s", ctypes.c_double),
                ("t_done_s", ctypes.c_double), ("nt", ctypes.c_int), ("nrec", ctypes.c_int)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError("q-vti_amd/libqvti.so is not built; run __graft_entry__.build() first")
        L = ctypes.CDLL(LIB_PATH)
        vp = ctypes.c_void_p
        i, f, d, ll = ctypes.c_int, ctypes.c_float, ctypes.c_double, ctypes.c_longlong
        sig = {
            "qvti_device_count": [vp], "qvti_create": [vp, ctypes.POINTER(Grid), i, i],
            "qvti_set_model": [vp] * 8, "qvti_forward": [vp, i, vp, vp, vp, vp, vp],
            "qvti_forward_batch": [vp, i, vp, vp, i, vp, vp, vp],
            "qvti_forward_acoustic_3D": [vp, i, i, i, i, i, i, i, i, i, i, i, f, f, f, f, f, f, f,
                                         vp, vp, vp, vp, vp, vp, vp, i, i, i, i, vp, vp, vp],
            "qvti_set_snapshot": [vp, i, i, vp, ll], "qvti_destroy": [vp],
            "qvti_begin": [vp, i, vp, vp, i], "qvti_step": [vp, i], "qvti_sync": [vp],
            "qvti_finish": [vp, vp, vp, vp], "qvti_get_state": [vp, i, i, vp], "qvti_steps_done": [vp],
            "qvti_enable_kernel_timing": [vp, i], "qvti_kernel_times": [vp, i, vp, i, vp, vp],
            "qvti_kernel_bytes": [vp, i, vp], "qvti_set_graphs": [vp, i], "qvti_set_probe": [vp, i, f],
            "qvti_fft3d": [i, i, i, i, vp], "qvti_fft3d_noalloc": [i, i, i, i, vp, vp], "qvti_batch_width": [vp, vp, i, vp, i, vp, vp],
            "qvti_forward_shots": [vp, i, vp, vp, i, vp, vp, vp],
            "qvti_set_media": [vp, i], "qvti_rtm": [vp, vp, i, vp, vp, vp, vp, vp, vp, vp, vp],
            "qvti_set_split": [vp, i, i], "qvti_get_split": [vp, vp], "qvti_tune_split": [vp, vp, vp],
            "qvti_save_wisdom": [ctypes.c_char_p], "qvti_load_wisdom": [ctypes.c_char_p], "qvti_deriv1d": [i, i, i, f, i, i, vp, vp],
            "qvti_ricker": [i, f, f, vp], "qvti_gamma_from_q": [vp, vp, ll],
            "qvti_reference_receivers": [i, i, i, i, vp, vp],
            "qvti_make_model": [i, ctypes.c_uint64, i, i, i, i, vp, vp, vp, vp, vp, vp, vp, vp],
            "qvti_server_create": [vp, i, ll, i, i], "qvti_server_submit": [vp, vp, vp],
            "qvti_server_poll": [vp, ctypes.c_uint64, vp, vp, vp, vp],
            "qvti_server_wait": [vp, ctypes.c_uint64, d], "qvti_server_cancel": [vp, ctypes.c_uint64],
            "qvti_server_destroy": [vp], "qvti_set_autotune": [vp, i], "qvti_tile_choice": [vp, i, vp, vp],
            "qvti_pad_model": [vp, i, i, i, i, vp], "qvti_load_model_raw": [ctypes.c_char_p, i, i, i, i, vp],
            "qvti_write_raw": [ctypes.c_char_p, vp, ll],
            "qvti_set_kspace": [vp, f, f, f, f],
            "qvti_checkpoint_bytes": [vp], "qvti_checkpoint": [vp, vp, ll], "qvti_restore": [vp, vp, ll],
            "qvti_set_snapshot_window": [vp, vp, vp],
        }
        for name, args in sig.items():
            if not hasattr(L, name):
                continue  # reported by the export test
            fn = getattr(L, name)
            fn.argtypes = args
            fn.restype = ctypes.c_int
        L.qvti_last_error.restype = ctypes.c_char_p
        L.qvti_device_bytes.argtypes = [vp]
        L.qvti_device_bytes.restype = ll
        L.qvti_estimate_bytes.argtypes = [ctypes.POINTER(Grid), i]
        L.qvti_estimate_bytes.restype = ll
        if hasattr(L, "qvti_checkpoint_bytes"):
            L.qvti_checkpoint_bytes.restype = ll
        if hasattr(L, "qvti_server_clock"):
            L.qvti_server_clock.argtypes = [vp]
            L.qvti_server_clock.restype = d
        _lib = L
    return _lib


def check(rc):
    if rc < 0:
        raise QvtiError(rc, lib().qvti_last_error().decode())
    return rc


def _p(a):
    return None if a is None else a.ctypes.data_as(ctypes.c_void_p)


def f32(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float32)


def i32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


def device_count():
    n = ctypes.c_int(0)
    rc = lib().qvti_device_count(ctypes.byref(n))
    return n.value if rc == 0 else 0


# ---------------------------------------------------------------- helpers
def ricker(nt, dt, f0):
    """Ricker wavelet of the reference host (viscoelas_3D.cpp:494-502)."""
    out = np.zeros(nt, np.float32)
    check(lib().qvti_ricker(nt, dt, f0, _p(out)))
    return out


def gamma_from_q(q):
    """gamma = atan(1/Q)/pi (viscoelas_3D.cpp:287-288)."""
    q = f32(q)
    g = np.zeros_like(q)
    check(lib().qvti_gamma_from_q(_p(q), _p(g), q.size))
    return g


def reference_receivers(pml, rnx_max, rny_max, dr):
    R = rnx_max * rny_max
    rx, ry = np.zeros(R, np.int32), np.zeros(R, np.int32)
    check(lib().qvti_reference_receivers(pml, rnx_max, rny_max, dr, _p(rx), _p(ry)))
    return rx, ry


def make_model(kind, seed, shape, params, nlay=1, vti=False):
    """Seeded synthetic model; returns dict(vp, vs, rho, qp, qs, gp, gs[, eps, delta])."""
    nx, ny, nz = shape
    out = {k: np.zeros(shape, np.float32) for k in ("vp", "vs", "rho", "qp", "qs")}
    e = np.zeros(shape, np.float32) if vti else None
    dl = np.zeros(shape, np.float32) if vti else None
    p = np.zeros(8, np.float64)
    p[:len(params)] = params
    check(lib().qvti_make_model(kind, seed, nx, ny, nz, nlay, _p(p), _p(out["vp"]), _p(out["vs"]),
                                _p(out["rho"]), _p(out["qp"]), _p(out["qs"]), _p(e), _p(dl)))
    out["gp"] = gamma_from_q(out["qp"])
    out["gs"] = gamma_from_q(out["qs"])
    if vti:
        out["eps"], out["delta"] = e, dl
    return out


def save_wisdom(path):
    check(lib().qvti_save_wisdom(os.fsencode(path)))


def load_wisdom(path):
    check(lib().qvti_load_wisdom(os.fsencode(path)))


def fft3d(x, sign=-1):
    """Unnormalised 3-D C2C on the engine's kernels (complex64 in/out)."""
    x = np.ascontiguousarray(x, np.complex64).copy()
    check(lib().qvti_fft3d(x.shape[0], x.shape[1], x.shape[2], sign, _p(x)))
    return x


def deriv1d(a, d, axis, sym):
    a = f32(a)
    out = np.zeros_like(a)
    check(lib().qvti_deriv1d(a.shape[0], a.shape[1], a.shape[2], d, axis, sym, _p(a), _p(out)))
    return out


def pad_model(interior, pml):
    """Edge-replicate an interior model (wx, wy, wz) into the total grid."""
    a = f32(interior)
    out = np.zeros(tuple(n + 2 * pml for n in a.shape), np.float32)
    check(lib().qvti_pad_model(_p(a), out.shape[0], out.shape[1], out.shape[2], pml, _p(out)))
    return out


def load_model_raw(path, shape, pml):
    """Interior raw FP32 file -> padded total-grid model of `shape`."""
    out = np.zeros(shape, np.float32)
    check(lib().qvti_load_model_raw(os.fsencode(path), shape[0], shape[1], shape[2], pml, _p(out)))
    return out


def write_raw(path, data):
    a = f32(data)
    check(lib().qvti_write_raw(os.fsencode(path), _p(a), a.size))


def make_shot(kind, src, rx, ry, rz):
    rx, ry, rz = i32(rx), i32(ry), i32(rz)
    ip = ctypes.POINTER(ctypes.c_int)
    s = Shot(kind, int(src[0]), int(src[1]), int(src[2]), len(rx), rx.ctypes.data_as(ip),
             ry.ctypes.data_as(ip), rz.ctypes.data_as(ip))
    s._keep = (rx, ry, rz)
    return s


MEDIA_VISCO, MEDIA_DISPERSION, MEDIA_LOSSLESS, MEDIA_QCOMP = 0, 1, 2, 3
IC_EXA, IC_SRCNORM, IC_XCORR = 1, 3, 4
FIELD_PRESSURE = 9


class RtmOpts(ctypes.Structure):
    _fields_ = [("condition", ctypes.c_int), ("field", ctypes.c_int), ("checkpoint_every", ctypes.c_int),
                ("eps", ctypes.c_float)]


def rtm(src, rcv, nt, rik, shot, recs, condition=IC_XCORR, field=FIELD_PRESSURE, checkpoint_every=500, eps=0.01):
    """Reverse-time migration of one shot (qvti_rtm): src, rcv are Engines of
    one grid with models set; recs = (vx, vy, vz) [nt][R] records of the
    shot's receivers, any None.  Returns (image, illumination) [nx][ny][nz]."""
    rik = f32(rik)
    rv = [None if r is None else f32(r) for r in recs]
    img = np.zeros(src.shape, np.float32)
    ill = np.zeros(src.shape, np.float32)
    o = RtmOpts(condition, field, checkpoint_every, eps)
    arr = (Shot * 1)(shot)
    check(lib().qvti_rtm(src._c, rcv._c, nt, _p(rik), ctypes.cast(arr, ctypes.c_void_p), *[_p(r) for r in rv],
                         ctypes.byref(o), _p(img), _p(ill)))
    return img, ill


class Engine:
    """One engine context (one grid, up to max_shots shots in flight)."""

    def __init__(self, nx, ny, nz, pml, dx, dy, dz, dt, f0, w0, vti=False, max_shots=1, device=0):
        self.grid = Grid(nx, ny, nz, pml, dx, dy, dz, dt, f0, w0, int(bool(vti)))
        self.shape = (nx, ny, nz)
        self._c = ctypes.c_void_p()
        check(lib().qvti_create(ctypes.byref(self._c), ctypes.byref(self.grid), max_shots, device))
        self.max_shots = max_shots
        self.nt = 0
        self._shots = None

    def close(self):
        if self._c:
            lib().qvti_destroy(self._c)
            self._c = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    @property
    def device_bytes(self):
        return lib().qvti_device_bytes(self._c)

    def set_model(self, vp, gp, vs, gs, rho, eps=None, delta=None):
        arrs = [f32(a) for a in (vp, gp, vs, gs, rho, eps, delta)]
        for a in arrs:
            if a is not None and a.shape != self.shape:
                raise ValueError("model array shape %s != grid %s" % (a.shape, self.shape))
        check(lib().qvti_set_model(self._c, *[_p(a) for a in arrs]))

    def set_model_dict(self, m):
        self.set_model(m["vp"], m["gp"], m["vs"], m["gs"], m["rho"], m.get("eps"), m.get("delta"))

    def set_split(self, axis, variant):
        check(lib().qvti_set_split(self._c, axis, variant))

    def split(self):
        v = np.zeros(3, np.int32)
        check(lib().qvti_get_split(self._c, _p(v)))
        return tuple(int(x) for x in v)

    def tune_split(self):
        """(chosen variants per axis, {(axis, variant): ms per step})."""
        v = np.zeros(3, np.int32)
        ms = np.full(6, np.nan)
        check(lib().qvti_tune_split(self._c, _p(v), _p(ms)))
        return tuple(int(x) for x in v), {(a // 2, a % 2): float(ms[a]) for a in range(6) if np.isfinite(ms[a])}

    def set_media(self, media):
        """Propagation medium (MEDIA_*) for the next set_model."""
        check(lib().qvti_set_media(self._c, media))

    def set_graphs(self, on):
        check(lib().qvti_set_graphs(self._c, int(on)))

    def set_autotune(self, on):
        check(lib().qvti_set_autotune(self._c, int(on)))

    def tile_choice(self):
        """Pencils per tile of each strided kernel class (after the first begin)."""
        names = ["tick", "vel_x", "vel_y", "vel_z_fwd", "y_fwd", "xpass", "y_inv", "z_inv", "str_x", "str_y"]
        out = np.zeros(32, np.int32)
        spl = np.zeros(32, np.int32)
        k = check(lib().qvti_tile_choice(self._c, 32, _p(out), _p(spl)))
        return {n: int(out[j]) for j, n in enumerate(names[:k]) if out[j] > 0}

    def tile_split(self):
        """Workgroups per tile (1 or the field count) of each strided kernel class."""
        names = ["tick", "vel_x", "vel_y", "vel_z_fwd", "y_fwd", "xpass", "y_inv", "z_inv", "str_x", "str_y"]
        out = np.zeros(32, np.int32)
        spl = np.zeros(32, np.int32)
        k = check(lib().qvti_tile_choice(self._c, 32, _p(out), _p(spl)))
        return {n: int(spl[j]) for j, n in enumerate(names[:k]) if out[j] > 0}

    def set_kspace(self, vref=0.0, fcut=0.0, vmax=0.0, taper_ratio=0.0):
        """Opt-in k-space correction sinc^2(vref |k| dt / 2) and radial Tukey
        wavenumber filter (cut-off 2 pi fcut / vmax) on the stress operators;
        all zero (the default) is the reference scheme."""
        check(lib().qvti_set_kspace(self._c, vref, fcut, vmax, taper_ratio))

    def steps_done(self):
        return lib().qvti_steps_done(self._c)

    def checkpoint(self):
        """Image (uint8 array) of the run in progress, after the queued steps."""
        n = lib().qvti_checkpoint_bytes(self._c)
        if n < 0:
            check(int(n))
        img = np.zeros(n, np.uint8)
        check(lib().qvti_checkpoint(self._c, _p(img), n))
        return img

    def restore(self, img):
        """Continue a run from an image; call right after the matching begin."""
        img = np.ascontiguousarray(img, np.uint8)
        check(lib().qvti_restore(self._c, _p(img), img.size))

    def set_probe(self, every, limit=1e30):
        check(lib().qvti_set_probe(self._c, every, limit))

    def set_snapshot(self, field, every, nt, window=None):
        """Snapshots of `field` every `every` steps over window = ((x0, y0, z0),
        (x1, y1, z1)) (default: the interior [pml, n - pml)^3); returns the
        [nsnap][wx][wy][wz] array they are written to by finish()."""
        g = self.grid
        if window is None:
            lo, hi = [g.pml] * 3, [g.nx - g.pml, g.ny - g.pml, g.nz - g.pml]
        else:
            lo, hi = list(window[0]), list(window[1])
        lo_a, hi_a = np.array(lo, np.int32), np.array(hi, np.int32)
        check(lib().qvti_set_snapshot_window(self._c, _p(lo_a), _p(hi_a)))
        ns = (nt + every - 1) // every if every else 0
        self._snap = np.zeros((max(ns, 1),) + tuple(int(h - l) for l, h in zip(lo, hi)), np.float32)
        check(lib().qvti_set_snapshot(self._c, field, every, _p(self._snap) if every else None,
                                      self._snap.size if every else 0))
        return self._snap

    def begin(self, nt, rik, shots):
        self._rik = f32(rik)
        arr = (Shot * len(shots))(*shots)
        self._shots = (arr, shots)
        check(lib().qvti_begin(self._c, nt, _p(self._rik), ctypes.cast(arr, ctypes.c_void_p), len(shots)))
        self.nt = nt

    def step(self, n):
        check(lib().qvti_step(self._c, n))

    def sync(self):
        check(lib().qvti_sync(self._c))

    def finish(self):
        shots = self._shots[1]
        R = sum(s.nrec for s in shots)
        recs = np.zeros((3, R * self.nt), np.float32)
        rc = lib().qvti_finish(self._c, _p(recs[0]), _p(recs[1]), _p(recs[2]))
        out, o = [], 0
        for s in shots:
            n = s.nrec * self.nt
            out.append(recs[:, o:o + n].reshape(3, self.nt, s.nrec))
            o += n
        check(rc)
        return out

    def forward(self, nt, rik, shot):
        R = shot.nrec
        rec = np.zeros((3, nt, R), np.float32)
        rik = f32(rik)
        arr = (Shot * 1)(shot)
        check(lib().qvti_forward(self._c, nt, _p(rik), ctypes.cast(arr, ctypes.c_void_p),
                                 _p(rec[0]), _p(rec[1]), _p(rec[2])))
        return rec

    def forward_batch(self, nt, rik, shots):
        R = sum(s.nrec for s in shots)
        recs = np.zeros((3, R * nt), np.float32)
        rik = f32(rik)
        arr = (Shot * len(shots))(*shots)
        check(lib().qvti_forward_batch(self._c, nt, _p(rik), ctypes.cast(arr, ctypes.c_void_p), len(shots),
                                       _p(recs[0]), _p(recs[1]), _p(recs[2])))
        out, o = [], 0
        for s in shots:
            n = s.nrec * nt
            out.append(recs[:, o:o + n].reshape(3, nt, s.nrec))
            o += n
        return out

    def batch_width(self, shots):
        """(shots per launch the library uses for these shots, {width: ms per shot-step tried})."""
        arr = (Shot * len(shots))(*shots)
        w = ctypes.c_int(0)
        tw = np.zeros(32, np.int32)
        tm = np.zeros(32, np.float64)
        k = check(lib().qvti_batch_width(self._c, ctypes.cast(arr, ctypes.c_void_p), len(shots), ctypes.byref(w), 32,
                                         _p(tw), _p(tm)))
        return w.value, {int(tw[j]): float(tm[j]) for j in range(k)}

    def forward_shots(self, nt, rik, shots):
        """All shots, in groups of the library's tuned width; list of [3][nt][R_s]."""
        R = sum(s.nrec for s in shots)
        recs = np.zeros((3, R * nt), np.float32)
        rik = f32(rik)
        arr = (Shot * len(shots))(*shots)
        check(lib().qvti_forward_shots(self._c, nt, _p(rik), ctypes.cast(arr, ctypes.c_void_p), len(shots),
                                       _p(recs[0]), _p(recs[1]), _p(recs[2])))
        out, o = [], 0
        for s in shots:
            n = s.nrec * nt
            out.append(recs[:, o:o + n].reshape(3, nt, s.nrec))
            o += n
        return out

    def state(self, shot=0):
        out = np.zeros((9,) + self.shape, np.float32)
        for f in range(9):
            check(lib().qvti_get_state(self._c, shot, f, _p(out[f])))
        return out

    def enable_kernel_timing(self, max_steps):
        check(lib().qvti_enable_kernel_timing(self._c, max_steps))

    def kernel_times(self):
        n = 32
        names = ctypes.create_string_buffer(n * 32)
        launches = np.zeros(n, np.int64)
        ms = np.zeros(n, np.float64)
        k = check(lib().qvti_kernel_times(self._c, n, names, 32, _p(launches), _p(ms)))
        byts = np.zeros(n, np.float64)
        check(lib().qvti_kernel_bytes(self._c, n, _p(byts)))
        res = {}
        for j in range(k):
            nm = names.raw[j * 32:(j + 1) * 32].split(b"\0")[0].decode()
            res[nm] = {"launches": int(launches[j]), "total_ms": float(ms[j]), "bytes_per_launch": float(byts[j])}
        return res


class Server:
    """Continuous-batching request server (qvti_server_*): requests of any
    supported grid are admitted by HBM budget and stepped concurrently."""

    def __init__(self, budget_bytes, max_active=4, steps_per_slice=25, device=0):
        self._s = ctypes.c_void_p()
        check(lib().qvti_server_create(ctypes.byref(self._s), device, int(budget_bytes), max_active,
                                       steps_per_slice))
        self._meta = {}

    def submit(self, grid, nt, rik, model, shot):
        """grid: Grid; model: dict vp, gp, vs, gs, rho (+ eps, delta when grid.vti)."""
        keep = [f32(rik)] + [f32(model.get(k)) for k in ("vp", "gp", "vs", "gs", "rho", "eps", "delta")]
        q = Request(grid, nt, *[_p(a) for a in keep[1:]], _p(keep[0]), shot)
        rid = ctypes.c_uint64()
        check(lib().qvti_server_submit(self._s, ctypes.byref(q), ctypes.byref(rid)))
        self._meta[rid.value] = (nt, shot.nrec)
        return rid.value

    def wait(self, rid, timeout_s=600.0):
        return lib().qvti_server_wait(self._s, rid, timeout_s)

    def cancel(self, rid):
        check(lib().qvti_server_cancel(self._s, rid))

    def info(self, rid):
        inf = ResultInfo()
        rc = lib().qvti_server_poll(self._s, rid, ctypes.byref(inf), None, None, None)
        return rc, inf

    def result(self, rid):
        """(status, info, records[3, nt, nrec] or None)."""
        nt, nrec = self._meta[rid]
        rec = np.zeros((3, nt, nrec), np.float32)
        inf = ResultInfo()
        rc = lib().qvti_server_poll(self._s, rid, ctypes.byref(inf), _p(rec[0]), _p(rec[1]), _p(rec[2]))
        return rc, inf, (rec if rc == 0 else None)

    def clock(self):
        return lib().qvti_server_clock(self._s)

    def close(self):
        if self._s:
            lib().qvti_server_destroy(self._s)
            self._s = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
