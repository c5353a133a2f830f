"""Physics pins of the oracle's time loop (CPU, no GPU).

The oracle restates the reference's loop (cuviscoelas_3D.cu:800-1101,
1255-1605).  The parts a misreading of that code could get wrong, and that
closed-form table tests do not reach, are pinned here by what the
mathematics fixes, independently of the oracle's own formulas:

* the Q operator wiring - which coefficient multiplies |k|^-1, 1, |k| and
  the backward time difference (cuviscoelas_3D.cu:195-205, 475-535, 879-898):
  the transfer function between two receivers in a homogeneous Q = 50 medium
  equals the plane-wave Green's function e^{ikr}(ikr - 1)/r^2 with k(w) from
  the scheme's constant-Q dispersion relation
      rho W^2 = k^2 [eta1/k + eta2 + eta3 k - i W e^{i w dt/2} (tau1/k + tau2)],
      W = (2/dt) sin(w dt/2)
  (low-loss phase speed c (1 + gamma (x - 1/x))^1/2, x = c k / w0, and
  Q^-1 ~ pi gamma);
* the D+/D- choice of the six shear derivatives, the buoyancy and shear
  averages and the half-cell C-PML profiles (cuviscoelas_3D.cu:1018-1101,
  902-1005): the lossless heterogeneous scheme, C-PML included, is exactly
  reciprocal for force sources - rho_B vz_B(fx at A) = rho_A vx_A(fz at B),
  rho the buoyancy average at the recorded component - to FP64 rounding;
* the S-wave path: SV speeds along x and z in a lossless VTI medium equal
  vs = sqrt(C44 / rho).

Each check is also run on the oracle with a plausible misreading planted
(oracle.SLIP_*: eta1 <-> eta3, eta signs reversed, tau terms dropped, D+ on
one shear derivative, node instead of half-cell C-PML profile in one memory
variable), and each planted misreading fails it.
"""
import functools

import numpy as np
import pytest

import oracle as O

PI_T = 3.1415926  # the reference's pi (cuviscoelas_3D.cu:15)


def _ricker(nt, dt, f0, delay=1.2):
    t = np.arange(nt) * dt
    a = (np.pi * f0 * (t - delay / f0)) ** 2
    return ((1 - 2 * a) * np.exp(-a)).astype(np.float32)


def _taper(nt, dt, t_arrival, half, ramp=0.02):
    """Window of half-width `half` s around an arrival, cosine ramps of `ramp` s."""
    t = np.arange(nt) * dt
    w = np.clip((half - np.abs(t - t_arrival)) / ramp, 0, 1)
    return 0.5 - 0.5 * np.cos(np.pi * w)


# ---------------------------------------------------------------- Q model
QCASE = dict(nx=96, ny=48, nz=48, pml=8, d=10.0, dt=1e-3, nt=200, vp=3000.0, vs=1730.0, rho=2000.0, Q=50.0,
             f0=30.0, src=(16, 24, 24), r=(14, 34))
Q_SLIPS = (0, O.SLIP_ETA_SWAP, O.SLIP_ETA_SIGN, O.SLIP_NO_TAU)
# propagation media run through the same check: (slip, media) keys
Q_MEDIA = (1, 3)  # dispersion only, Q-compensating


def _concurrent(fn, args, threads, workers=8):
    """fn(*a, nthreads) for every a in args on up to `workers` concurrent host
    threads (the oracle releases the GIL), `threads` OpenMP threads each:
    small grids run faster as many one-thread runs than one many-thread run."""
    from concurrent.futures import ThreadPoolExecutor

    with ThreadPoolExecutor(min(workers, len(args))) as ex:
        return list(ex.map(lambda a: fn(*a, threads), args))


@functools.lru_cache(maxsize=None)
def _q_runs():
    keys = [(s, 0) for s in Q_SLIPS] + [(0, m) for m in Q_MEDIA]
    return dict(zip(keys, _concurrent(_q_one, keys, 2)))


def _q_run(slip, media=0):
    return _q_runs()[(slip, media)]


def _q_one(slip, media, nthreads):
    c = QCASE
    shape = (c["nx"], c["ny"], c["nz"])
    g = np.arctan(1 / c["Q"]) / np.pi
    m = dict(vp=np.full(shape, c["vp"], np.float32), vs=np.full(shape, c["vs"], np.float32),
             rho=np.full(shape, c["rho"], np.float32), gp=np.full(shape, g, np.float32),
             gs=np.full(shape, g, np.float32))
    sx, sy, sz = c["src"]
    shot = dict(nx=c["nx"], ny=c["ny"], nz=c["nz"], pml=c["pml"], nt=c["nt"], dx=c["d"], dy=c["d"], dz=c["d"],
                dt=c["dt"], f0=c["f0"], w0=float(np.float32(2 * PI_T * c["f0"])), src_kind=O.SRC_EXPLOSIVE,
                sx=sx, sy=sy, sz=sz, rx=np.array([sx + c["r"][0], sx + c["r"][1]]), ry=np.full(2, sy),
                rz=np.full(2, sz), rik=_ricker(c["nt"], c["dt"], c["f0"]))
    rec = O.run(shot, m, "f32", nthreads=nthreads, slip=slip, media=media)["records"][0]
    return rec.astype(np.float64), shot["w0"], g


def _q_wavenumber(omega, rho, vp, g, w0, dt, tau_scale=1.0):
    """Complex wavenumber of the scheme's P wave at real frequencies: the root
    near W/c of eta3 k^3 + (eta2 - i W' tau2) k^2 + (eta1 - i W' tau1) k -
    rho W^2 = 0, W = (2/dt) sin(w dt/2), W' = W e^{i w dt/2}, with the
    coefficients of the reference's constant-Q model (c = v cos(pi g / 2)),
    the loss terms scaled by tau_scale (0: dispersion only, -1: the
    Q-compensating medium)."""
    c = vp * np.cos(g * np.pi / 2)
    e1, e2, e3 = -rho * c * g * w0, rho * c * c, rho * c ** 3 * g / w0
    t1 = tau_scale * rho * c * g * np.pi
    t2 = tau_scale * rho * c * c * g * g * np.pi / w0
    out = []
    for w in omega:
        W = 2 / dt * np.sin(w * dt / 2)
        Wp = W * np.exp(1j * w * dt / 2)
        roots = np.roots([e3, e2 - 1j * Wp * t2, e1 - 1j * Wp * t1, -rho * W * W])
        out.append(roots[np.argmin(np.abs(roots - W / c))])
    return np.array(out)


def _q_misfit(slip, media=0, tau_scale=1.0):
    """Relative misfit of the measured receiver-to-receiver transfer function
    against the plane-wave Green's function over 15-55 Hz."""
    c = QCASE
    rec, w0, g = _q_run(slip, media)
    if not np.all(np.isfinite(rec)):
        return np.inf, np.inf, None
    # vx sits half a cell behind its cell along x (D- derivative of pxx)
    r1, r2 = (c["r"][0] - 0.5) * c["d"], (c["r"][1] - 0.5) * c["d"]
    nf = 4096
    spec = []
    for j, r in enumerate((r1, r2)):
        win = _taper(c["nt"], c["dt"], 1.2 / c["f0"] + r / c["vp"], 0.04)
        spec.append(np.fft.rfft(rec[:, j] * win, nf))
    f = np.fft.rfftfreq(nf, c["dt"])
    band = (f >= 15) & (f <= 55)
    H = np.conj(spec[1][band] / spec[0][band])  # e^{-i w t} convention
    k = _q_wavenumber(2 * np.pi * f[band], c["rho"], c["vp"], g, w0, c["dt"], tau_scale)
    Hp = np.exp(1j * k * (r2 - r1)) * (1j * k * r2 - 1) * r1 ** 2 / ((1j * k * r1 - 1) * r2 ** 2)
    err = np.abs(H - Hp) / np.abs(Hp)
    return float(err.max()), float(np.sqrt(np.mean(err ** 2))), (f[band], k)


def test_constant_q_dispersion_and_loss():
    """Q = 50 (gamma = 0.00637): the measured transfer function between
    receivers 135 m and 335 m from an explosive source matches the scheme's
    constant-Q plane-wave Green's function within 3 % at every frequency of
    15-55 Hz (phase speed rising with frequency as 1 + gamma (x - 1/x), loss
    ~ pi gamma per radian)."""
    mx, rms, (f, k) = _q_misfit(0)
    c = QCASE
    for fi in (20.0, 30.0, 40.0, 50.0):
        j = int(np.argmin(np.abs(f - fi)))
        cp = 2 * np.pi * f[j] / k[j].real
        print("%.0f Hz: phase speed %.1f m/s, Q %.1f" % (f[j], cp, k[j].real / (2 * k[j].imag)))
    print("transfer-function misfit: max %.4f rms %.4f" % (mx, rms))
    assert mx < 0.03 and rms < 0.015, (mx, rms)
    # the model is dispersive and lossy at the level the check resolves
    assert 2 * np.pi * f[-1] / k[-1].real - 2 * np.pi * f[0] / k[0].real > 0.005 * c["vp"]


@pytest.mark.parametrize("slip", [O.SLIP_ETA_SWAP, O.SLIP_ETA_SIGN, O.SLIP_NO_TAU])
def test_constant_q_check_fails_on_planted_slips(slip):
    """eta1 <-> eta3, reversed eta signs (dispersion term 1 - gamma (x - 1/x))
    and dropped tau terms (no loss) each miss the Green's function by more
    than 5 % rms."""
    mx, rms, _ = _q_misfit(slip)
    print("slip %d: misfit max %.3f rms %.3f" % (slip, mx, rms))
    assert rms > 0.05


@pytest.mark.parametrize("media,tau_scale", [(1, 0.0), (3, -1.0)])
def test_dispersion_only_and_q_compensating_media(media, tau_scale):
    """The oracle's media used by reverse-time migration (the reference's
    media_type 2 and 4, viscoelas_3D.cpp:57-60): dispersion-only keeps the
    constant-Q phase speed with no loss, the Q-compensating medium turns the
    loss into a gain of the same size - each matches the Green's function of
    its own dispersion relation (tau scaled by 0 or -1) within 3 %, and
    misses the constant-Q one."""
    mx, rms, (f, k) = _q_misfit(0, media, tau_scale)
    other, _, _ = _q_misfit(0, media, 1.0)
    print("media %d: misfit max %.4f rms %.4f (against the lossy model: max %.3f); Im k at 40 Hz %.2e" % (
        media, mx, rms, other, k[np.argmin(np.abs(f - 40))].imag))
    assert mx < 0.03 and rms < 0.015
    assert other > 0.1


def test_injected_sources_enter_like_the_shots_own_source():
    """An extra point source carrying the wavelet at the source cell, with the
    shot's own wavelet silent, gives the same run bit for bit (both are added
    after the stress update and enter the next step's V - V_prev); the
    lossless medium equals gamma = 0 bit for bit."""
    from cases import small_case

    shot, model = small_case(16, 20, 24, 3, 25, O.SRC_FZ, True)
    a = O.run(shot, model, "f64", state=True)
    silent = dict(shot, rik=np.zeros_like(shot["rik"]))
    tr = np.asarray(shot["rik"], np.float32)[:, None]
    inj = dict(x=[shot["sx"]], y=[shot["sy"]], z=[shot["sz"]], vx=None, vy=None, vz=tr)
    b = O.run(silent, model, "f64", state=True, inject=inj)
    assert np.array_equal(a["state"], b["state"]) and np.array_equal(a["records"], b["records"])
    c = O.run(shot, model, "f64", state=True, media=2)
    d = O.run(shot, dict(model, gp=np.zeros_like(model["gp"]), gs=np.zeros_like(model["gs"])), "f64", state=True)
    assert np.array_equal(c["state"], d["state"])


# ----------------------------------------------------------- reciprocity
RCASE = dict(nx=32, ny=28, nz=30, pml=5, d=10.0, dt=1e-3, nt=80, f0=45.0, A=(13, 12, 16), B=(19, 16, 11))


def _recip_model(vti):
    c = RCASE
    shape = (c["nx"], c["ny"], c["nz"])
    rng = np.random.default_rng(5)
    vp = (2500 + 1000 * rng.random(shape)).astype(np.float32)
    m = dict(vp=vp, vs=(vp / 1.8).astype(np.float32), rho=(1800 + 600 * rng.random(shape)).astype(np.float32),
             gp=np.zeros(shape, np.float32), gs=np.zeros(shape, np.float32))
    if vti:
        m["eps"] = (0.2 * rng.random(shape)).astype(np.float32)
        m["delta"] = (m["eps"] * rng.random(shape)).astype(np.float32)
    return m


R_CASES = ((False, 0), (True, 0), (False, O.SLIP_SHEAR_DPLUS), (False, O.SLIP_PML_NODE_V),
           (False, O.SLIP_PML_NODE_S))


@functools.lru_cache(maxsize=None)
def _recip_all():
    runs = _concurrent(_recip_run, [(v, s, w) for v, s in R_CASES for w in (0, 1)], 1)
    return {case: _recip_compare(case[0], runs[2 * j], runs[2 * j + 1]) for j, case in enumerate(R_CASES)}


def _recip_residual(vti, slip):
    return _recip_all()[(vti, slip)]


def _recip_run(vti, slip, which, nthreads):
    """which 0: x-force at A, vz recorded at B; 1: z-force at B, vx recorded at A."""
    c = RCASE
    m = _recip_model(vti)
    A, B = c["A"], c["B"]
    src, rec, kind, comp = (A, B, O.SRC_FX, 2) if which == 0 else (B, A, O.SRC_FZ, 0)
    shot = dict(nx=c["nx"], ny=c["ny"], nz=c["nz"], pml=c["pml"], nt=c["nt"], dx=c["d"], dy=c["d"], dz=c["d"],
                dt=c["dt"], f0=c["f0"], w0=float(np.float32(2 * PI_T * c["f0"])), rik=_ricker(c["nt"], c["dt"], c["f0"]),
                src_kind=kind, sx=src[0], sy=src[1], sz=src[2], rx=np.array([rec[0]]), ry=np.array([rec[1]]),
                rz=np.array([rec[2]]))
    return O.run(shot, m, "f64", nthreads=nthreads, slip=slip)["records"][comp][:, 0]


def _recip_compare(vti, vz_b, vx_a):
    """|| rho_B vz_B(fx at A) - rho_A vx_A(fz at B) || / || rho_A vx_A(fz at B) ||
    on a lossless heterogeneous model (per-cell random vp, vs, rho, eps,
    delta), rho the buoyancy averages of the recorded components."""
    c = RCASE
    A, B = c["A"], c["B"]
    rho = _recip_model(vti)["rho"].astype(np.float64)
    rho_b = 0.5 * (rho[B] + rho[B[0], B[1], B[2] + 1])  # vz: average toward +z
    rho_a = 0.5 * (rho[A] + rho[A[0] + 1, A[1], A[2]])  # vx: average toward +x
    u, v = vz_b * rho_b, vx_a * rho_a
    if not (np.all(np.isfinite(u)) and np.all(np.isfinite(v))):
        return np.inf
    return float(np.linalg.norm(u - v) / np.linalg.norm(v))


@pytest.mark.parametrize("vti", [False, True])
def test_lossless_scheme_is_reciprocal(vti):
    """The staggered pairs (D- on an own-axis normal stress in the velocity
    step against D+ on the own-axis velocity derivative in the stress step,
    D+ against D- for the shear pairs) are anti-adjoint and every material
    and C-PML factor is diagonal on the grid its equation lives on, so the
    lossless scheme - C-PML, 8-corner shear averages and last-plane skip
    included - is reciprocal to rounding.  The waves reach the C-PML within
    the window."""
    r = _recip_residual(vti, 0)
    print("reciprocity residual (%s): %.2e" % ("VTI" if vti else "isotropic", r))
    assert r < 1e-12


@pytest.mark.parametrize("slip", [O.SLIP_SHEAR_DPLUS, O.SLIP_PML_NODE_V, O.SLIP_PML_NODE_S])
def test_reciprocity_fails_on_planted_slips(slip):
    """D+ on the z-derivative of vx in the pxz update, or a node instead of a
    half-cell C-PML profile in one memory variable, breaks reciprocity by
    orders of magnitude above rounding."""
    r = _recip_residual(False, slip)
    print("slip %d: reciprocity residual %.2e" % (slip, r))
    assert r > 1e-7


# --------------------------------------------------------------- S waves
SCASE = dict(nx=48, ny=20, nz=48, pml=8, d=10.0, dt=1e-3, nt=180, vp=3000.0, vs=1500.0, f0=30.0, r=(8, 16))


def _lag(a, b, dt):
    """Delay of trace b relative to a by cross-correlation, parabolic refinement."""
    cc = np.correlate(b, a, mode="full")
    k = int(np.argmax(cc))
    y0, y1, y2 = cc[k - 1], cc[k], cc[k + 1]
    return (k - (len(a) - 1) + 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)) * dt


S_CASES = ((0, 0), (2, 0), (2, O.SLIP_SHEAR_DPLUS))


@functools.lru_cache(maxsize=None)
def _s_all():
    return dict(zip(S_CASES, _concurrent(_s_one, S_CASES, 2)))


def _s_speed(axis, slip):
    return _s_all()[(axis, slip)]


def _s_one(axis, slip, nthreads):
    """SV speed / vs from the moveout between receivers 80 m and 160 m from a
    force source in a lossless homogeneous VTI medium (eps 0.2, delta 0.1):
    axis 0 - z-force, wave along x, vz recorded; axis 2 - x-force, wave along
    z, vx recorded.  Traces are windowed around the S arrivals."""
    c = SCASE
    shape = (c["nx"], c["ny"], c["nz"])
    m = dict(vp=np.full(shape, c["vp"], np.float32), vs=np.full(shape, c["vs"], np.float32),
             rho=np.full(shape, 2000, np.float32), gp=np.full(shape, 1e-6, np.float32),
             gs=np.full(shape, 1e-6, np.float32), eps=np.full(shape, 0.2, np.float32),
             delta=np.full(shape, 0.1, np.float32))
    r1, r2 = c["r"]
    if axis == 0:
        src, rx, rz, kind, comp = (14, 10, 24), [14 + r1, 14 + r2], [24, 24], O.SRC_FZ, 2
    else:
        src, rx, rz, kind, comp = (24, 10, 14), [24, 24], [14 + r1, 14 + r2], O.SRC_FX, 0
    shot = dict(nx=c["nx"], ny=c["ny"], nz=c["nz"], pml=c["pml"], nt=c["nt"], dx=c["d"], dy=c["d"], dz=c["d"],
                dt=c["dt"], f0=c["f0"], w0=float(np.float32(2 * PI_T * c["f0"])), src_kind=kind, sx=src[0],
                sy=src[1], sz=src[2], rx=np.array(rx), ry=np.full(2, 10), rz=np.array(rz),
                rik=_ricker(c["nt"], c["dt"], c["f0"]))
    rec = O.run(shot, m, "f32", nthreads=nthreads, slip=slip)["records"][comp].astype(np.float64)
    if not np.all(np.isfinite(rec)) or np.abs(rec).max() > 1e6:
        return np.inf
    tr = [rec[:, j] * _taper(c["nt"], c["dt"], 1.2 / c["f0"] + r * c["d"] / c["vs"], 0.045)
          for j, r in enumerate((r1, r2))]
    return (r2 - r1) * c["d"] / _lag(tr[0], tr[1], c["dt"]) / c["vs"]


@pytest.mark.parametrize("axis", [0, 2])
def test_s_wave_speed_equals_vs(axis):
    s = _s_speed(axis, 0)
    print("SV speed along %s: %.4f vs" % ("xz"[axis // 2], s))
    assert abs(s - 1) < 0.01


def test_s_wave_check_fails_on_shear_slip():
    """D+ on the z-derivative of vx in the pxz update: the S wave along z is
    no longer a stable wave at vs."""
    s = _s_speed(2, O.SLIP_SHEAR_DPLUS)
    print("slip: SV speed along z %.4f vs" % s)
    assert not abs(s - 1) < 0.01
