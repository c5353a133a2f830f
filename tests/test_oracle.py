"""CPU tests that pin the oracle (oracle/qvti_oracle.c) before it is trusted.

The reference ships no tests, fixtures or golden records (its model blob is
missing), so every part of the oracle is pinned here by something independent
of its own formulas:
  * the FFT by numpy/pocketfft FP64 vectors (tests/golden/fft_vectors.npz) and
    by a brute-force DFT;
  * the D+/D- tables, |K| tables, C-PML profiles, Q coefficients and Ricker by
    closed forms and their identities;
  * the time loop by exact linearity, the VTI -> isotropic identity, the
    factorised restatement of tests/factorised.py (a different algorithm that
    shares only the 1-D tables) and a travel-time check of the VTI physics.
"""
import os

import numpy as np
import pytest

import oracle as O
from cases import rel_l2, small_case

HERE = os.path.dirname(os.path.abspath(__file__))
PI_T = 3.1415926  # the reference's truncated pi


@pytest.fixture(scope="module")
def fft_golden():
    return np.load(os.path.join(HERE, "golden", "fft_vectors.npz"))


def test_fft1d_matches_pocketfft_golden(fft_golden):
    sizes = sorted(int(k[1:]) for k in fft_golden.files if k.startswith("x") and k[1:].isdigit())
    assert len(sizes) >= 30
    for n in sizes:
        x, X = fft_golden["x%d" % n], fft_golden["X%d" % n]
        scale = np.abs(X).max()
        assert np.abs(O.fft1d(x, -1, "f64") - X).max() <= 1e-13 * scale * max(1, np.log2(n)), n
        assert np.abs(O.fft1d(x, -1, "f32") - X).max() <= 1e-6 * scale, n
        # inverse: conj(FFT(conj(X))) = n * x
        assert np.abs(O.fft1d(X, +1, "f64") - n * x).max() <= 1e-13 * n * np.abs(x).max() * max(1, np.log2(n)), n


def test_fft3d_matches_pocketfft_golden(fft_golden):
    x, X = fft_golden["x3d"], fft_golden["X3d"]
    assert np.abs(O.fft3d(x, -1, "f64") - X).max() <= 1e-12 * np.abs(X).max()
    assert np.abs(O.fft3d(X, +1, "f64") - x.size * x).max() <= 1e-12 * np.abs(X).max()


@pytest.mark.parametrize("n", [1, 2, 3, 5, 6, 7, 12, 14, 20, 25, 30])
def test_fft_matches_bruteforce_dft(n):
    rng = np.random.default_rng(n)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    j = np.arange(n)
    W = np.exp(-2j * np.pi * np.outer(j, j) / n)
    ref = W @ x
    assert np.abs(O.fft1d(x, -1, "f64") - ref).max() <= 1e-12 * max(1.0, np.abs(ref).max())
    assert np.abs(O.fft1d(x, +1, "f64") - W.conj() @ x).max() <= 1e-12 * max(1.0, np.abs(ref).max())


def test_fft_delta_constant_and_parseval():
    n = (6, 10, 8)
    d = np.zeros(n, complex)
    d[0, 0, 0] = 1
    assert np.allclose(O.fft3d(d), np.ones(n), atol=1e-15)
    c = np.ones(n, complex)
    F = O.fft3d(c)
    assert abs(F[0, 0, 0] - c.size) < 1e-12 and np.abs(F.ravel()[1:]).max() < 1e-12
    rng = np.random.default_rng(3)
    x = rng.standard_normal(n) + 1j * rng.standard_normal(n)
    X = O.fft3d(x)
    assert abs(np.sum(np.abs(X) ** 2) / x.size - np.sum(np.abs(x) ** 2)) < 1e-10 * np.sum(np.abs(x) ** 2)


@pytest.mark.parametrize("n,d", [(8, 10.0), (15, 12.5), (64, 10.0), (100, 7.0)])
def test_axis_tables_closed_form(n, d):
    """cuviscoelas_3D.cu:44-149: folded |K| and D+- = iK exp(+-iK d/2), K signed."""
    kf, dp, dm = O.axis_tables(n, d, "f64")
    m = np.arange(n)
    fold = np.where(m <= n // 2, m, n - m)
    K = 2 * PI_T * fold / (n * d) * np.where(m <= n // 2, 1, -1)
    assert np.allclose(kf, np.abs(K), rtol=1e-14, atol=0)
    assert np.allclose(dp, 1j * K * np.exp(1j * K * d / 2), rtol=1e-13, atol=1e-15)
    assert np.allclose(dm, 1j * K * np.exp(-1j * K * d / 2), rtol=1e-13, atol=1e-15)
    # conjugate symmetry away from Nyquist (what makes C2C + .x real-to-real)
    for i in range(1, (n + 1) // 2):
        assert dp[n - i] == np.conj(dp[i]) and dm[n - i] == np.conj(dm[i])


def test_shift_derivative_of_sine():
    """D+ differentiates at x + d/2, D- at x - d/2 (band-limited sine)."""
    n, d, m0 = 48, 10.0, 5
    j = np.arange(n)
    f = np.sin(2 * np.pi * m0 * j / n)
    _, dp, dm = O.axis_tables(n, d, "f64")
    Kt = 2 * PI_T * m0 / (n * d)
    for D, s in ((dp, 1), (dm, -1)):
        got = np.fft.ifft(np.fft.fft(f) * D).real
        ref = Kt * np.cos(2 * np.pi * m0 * j / n + s * PI_T * m0 / n)
        assert np.abs(got - ref).max() < 1e-13


@pytest.mark.parametrize("n,pml,d,f0", [(30, 10, 10.0, 25.0), (64, 20, 12.5, 20.0), (12, 2, 5.0, 40.0)])
def test_pml_profile_closed_form(n, pml, d, f0):
    """cuviscoelas_3D.cu:231-307 with vmax forced to 5000 and kappa = 1."""
    a, b = O.pml_profile(n, pml, d, f0, "f64")
    R = float(np.float32(1e-2))
    amax = PI_T * f0
    om_max = 4 * 5000.0 * np.log(1.0 / R) / (2 * pml * d)
    i = np.arange(n)
    j = np.where(i < pml, i, np.where(i > n - 1 - pml, n - 1 - i, -1))
    alpha = np.where(j >= 0, amax * (np.maximum(j, 0) / (pml - 1)) ** 2, amax)
    om = np.where(j >= 0, om_max * ((pml - 1 - np.maximum(j, 0)) / pml) ** 3, 0.0)
    assert np.allclose(a, alpha + om, rtol=1e-12)
    assert np.allclose(b, om, rtol=1e-12, atol=0)
    assert b[pml - 1] == 0 and b[n - pml] == 0 and np.all(b[pml:n - pml] == 0)
    a32, b32 = O.pml_profile(n, pml, d, f0, "f32")
    assert np.allclose(a32, a, rtol=2e-6) and np.allclose(b32, b, rtol=2e-6, atol=1e-30)


def test_q_coefficient_identities():
    """The 5-term operator is rho c^2 [1 + gamma (x - 1/x)] and
    rho c^2 (pi gamma / w0) (1/x + gamma), x = c|k|/w0 (cuviscoelas_3D.cu:195-205)."""
    rng = np.random.default_rng(5)
    w0 = float(np.float32(2 * PI_T * 25))
    for _ in range(20):
        rho, v, Q = 1500 + 1500 * rng.random(), 1000 + 4000 * rng.random(), 10 + 190 * rng.random()
        g = np.arctan(1 / Q) / PI_T
        e1, e2, e3, t1, t2 = O.q_coeffs(rho, v, g, w0)
        c = v * np.cos(g * PI_T / 2)
        for k in (1e-3, 1e-2, 0.1):
            x = c * k / w0
            assert np.isclose(e1 / k + e2 + e3 * k, rho * c * c * (1 + g * (x - 1 / x)), rtol=1e-12)
            assert np.isclose(t1 / k + t2, rho * c * c * (PI_T * g / w0) * (1 / x + g), rtol=1e-12)
    assert np.allclose(O.q_coeffs(2000, 3000, 0.0, w0), [0, 2000 * 3000.0 ** 2, 0, 0, 0], rtol=1e-15, atol=0)


def test_oracle_is_exactly_linear_in_the_source():
    shot, model = small_case(16, 12, 10, 3, 12, O.SRC_FZ, False)
    a = O.run(shot, model, "f64")["records"]
    shot2 = dict(shot, rik=(2 * shot["rik"]).astype(np.float32))
    b = O.run(shot2, model, "f64")["records"]
    assert np.array_equal(2 * a, b)
    a32 = O.run(shot, model, "f32")["records"]
    b32 = O.run(shot2, model, "f32")["records"]
    assert np.array_equal(2 * a32, b32)


def test_vti_with_zero_thomsen_equals_isotropic():
    shot, model = small_case(12, 10, 16, 3, 10, O.SRC_EXPLOSIVE, False)
    iso = O.run(shot, model, "f64", state=True)
    zero = np.zeros_like(model["vp"])
    vti = O.run(shot, dict(model, eps=zero, delta=zero), "f64", state=True)
    assert np.array_equal(iso["state"], vti["state"])
    iso32 = O.run(shot, model, "f32", state=True)
    vti32 = O.run(shot, dict(model, eps=zero, delta=zero), "f32", state=True)
    assert np.array_equal(iso32["state"], vti32["state"])


@pytest.mark.parametrize("case", [(8, 10, 12, 3, 12, O.SRC_FX, False), (10, 8, 12, 2, 10, O.SRC_EXPLOSIVE, True),
                                  (12, 10, 8, 3, 10, O.SRC_FZ, False)])
def test_literal_oracle_equals_factorised_scheme_fp64(case):
    """63 3-D transforms per step (oracle) vs 1-D derivatives of {A1 v, v, A3 v,
    A1 dv, dv} (the engine's algebra), both FP64: equal to rounding."""
    import factorised

    shot, model = small_case(*case)
    ref = O.run(shot, model, "f64", state=True)
    rec, st = factorised.run(shot, model, case[6])
    assert rel_l2(st, ref["state"]) < 1e-12
    assert rel_l2(rec, ref["records"]) < 1e-10


def test_fp32_oracle_close_to_fp64():
    shot, model = small_case(16, 20, 24, 4, 30, O.SRC_EXPLOSIVE, True)
    a = O.run(shot, model, "f32", state=True)
    b = O.run(shot, model, "f64", state=True)
    assert rel_l2(a["state"], b["state"]) < 2e-6


def _lag(a, b, dt):
    """Delay of trace b relative to a by cross-correlation, parabolic refinement."""
    cc = np.correlate(b, a, mode="full")
    k = int(np.argmax(cc))
    y0, y1, y2 = cc[k - 1], cc[k], cc[k + 1]
    frac = 0.5 * (y0 - y2) / (y0 - 2 * y1 + y2)
    return (k - (len(a) - 1) + frac) * dt


def test_vti_horizontal_and_vertical_p_speeds():
    """Lossless-limit physics of the VTI extension: in a homogeneous VTI medium
    the qP wave travels at vp sqrt(1 + 2 eps) along x and at vp along z
    (moveout between two receivers 140 m and 300 m from an explosive source on
    each axis; the isotropic run of the same geometry measures 1.008 vp)."""
    nx, ny, nz, pml, d, dt, nt = 96, 32, 96, 8, 10.0, 1e-3, 160
    vp, vs, eps, delta = 3000.0, 1500.0, 0.2, 0.1
    shape = (nx, ny, nz)
    m = dict(vp=np.full(shape, vp, np.float32), vs=np.full(shape, vs, np.float32),
             rho=np.full(shape, 2000, np.float32), gp=np.full(shape, 1e-6, np.float32),
             gs=np.full(shape, 1e-6, np.float32), eps=np.full(shape, eps, np.float32),
             delta=np.full(shape, delta, np.float32))
    cx, cy, cz, r1, r2, f0 = 48, 16, 48, 14, 30, 30.0
    t = np.arange(nt) * dt
    a = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    rik = ((1 - 2 * a) * np.exp(-a)).astype(np.float32)
    shot = dict(nx=nx, ny=ny, nz=nz, pml=pml, nt=nt, dx=d, dy=d, dz=d, dt=dt, f0=f0,
                w0=float(np.float32(2 * PI_T * f0)), src_kind=O.SRC_EXPLOSIVE, sx=cx, sy=cy, sz=cz,
                rx=np.array([cx + r1, cx + r2, cx, cx]), ry=np.full(4, cy), rz=np.array([cz, cz, cz + r1, cz + r2]),
                rik=rik)
    rec = O.run(shot, m, "f32")["records"]
    vx = (r2 - r1) * d / _lag(rec[0, :, 0], rec[0, :, 1], dt)
    vz = (r2 - r1) * d / _lag(rec[2, :, 2], rec[2, :, 3], dt)
    assert abs(vz / vp - 1) < 0.02, vz
    assert abs((vx / vz) / np.sqrt(1 + 2 * eps) - 1) < 0.01, (vx, vz)


def _qp_group_speed_45(vp, vs, eps, delta):
    """Exact qP group speed at 45 degrees from the symmetry axis of a VTI
    medium: the exact phase velocity in Thomsen's parameters (Tsvankin,
    Seismic Signatures and Analysis of Reflection Data in Anisotropic Media,
    eq. 1.56) and the group velocity V = v sqrt(1 + (v'/v)^2) at the phase
    angle whose group angle tan(psi) = (tan th + v'/v) / (1 - tan th v'/v) is 45
    degrees."""
    f = 1 - (vs / vp) ** 2
    th = np.linspace(0, np.pi / 2, 200001)
    s2 = np.sin(th) ** 2
    v = vp * np.sqrt(1 + eps * s2 - f / 2 + f / 2 * np.sqrt((1 + 2 * eps * s2 / f) ** 2
                                                             - 2 * (eps - delta) * np.sin(2 * th) ** 2 / f))
    dv = np.gradient(v, th)
    psi = np.arctan2(np.tan(th) + dv / v, 1 - np.tan(th) * dv / v)
    return (v * np.sqrt(1 + (dv / v) ** 2))[np.argmin(np.abs(psi - np.pi / 4))]


def test_vti_diagonal_p_speed_carries_delta():
    """delta enters the scheme only through C13 = sqrt((C33-C44)(C33(1+2 delta)-C44)) - C44,
    which the axis speeds above do not see.  Along the 45-degree diagonal of the
    x-z plane the qP moveout (receivers 141 m and 297 m from an explosive
    source) matches the exact group speed of the textbook VTI dispersion
    relation within 2 % for delta = +/-0.1 (eps = 0.2), and the change between
    the two deltas (4.7 % of vp) is reproduced within 25 %."""
    nx, ny, nz, pml, d, dt, nt = 96, 32, 96, 8, 10.0, 1e-3, 160
    vp, vs, eps = 3000.0, 1500.0, 0.2
    shape = (nx, ny, nz)
    cx, cy, cz, r1, r2, f0 = 48, 16, 48, 10, 21, 30.0
    t = np.arange(nt) * dt
    a = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    rik = ((1 - 2 * a) * np.exp(-a)).astype(np.float32)
    shot = dict(nx=nx, ny=ny, nz=nz, pml=pml, nt=nt, dx=d, dy=d, dz=d, dt=dt, f0=f0,
                w0=float(np.float32(2 * PI_T * f0)), src_kind=O.SRC_EXPLOSIVE, sx=cx, sy=cy, sz=cz,
                rx=np.array([cx + r1, cx + r2]), ry=np.full(2, cy), rz=np.array([cz + r1, cz + r2]), rik=rik)
    meas, exact = {}, {}
    for delta in (0.1, -0.1):
        m = dict(vp=np.full(shape, vp, np.float32), vs=np.full(shape, vs, np.float32),
                 rho=np.full(shape, 2000, np.float32), gp=np.full(shape, 1e-6, np.float32),
                 gs=np.full(shape, 1e-6, np.float32), eps=np.full(shape, eps, np.float32),
                 delta=np.full(shape, delta, np.float32))
        rec = O.run(shot, m, "f32")["records"]
        u = (rec[0] + rec[2]) / np.sqrt(2)  # velocity along the diagonal
        meas[delta] = (r2 - r1) * d * np.sqrt(2) / _lag(u[:, 0], u[:, 1], dt) / vp
        exact[delta] = _qp_group_speed_45(vp, vs, eps, delta) / vp
        assert abs(meas[delta] / exact[delta] - 1) < 0.02, (delta, meas[delta], exact[delta])
    dm, de = meas[0.1] - meas[-0.1], exact[0.1] - exact[-0.1]
    assert de > 0.04 and abs(dm / de - 1) < 0.25, (dm, de)
