"""CPU tests that pin the oracle's opt-in k-space mode (the k-space time-stepping
correction sinc^2(vref |k| dt / 2) and the radial Tukey wavenumber filter
multiplied into the stress operators A1, A2, A3).  The reference has both
switched off (sinc2nd = 1 at cuviscoelas_3D.cu:183-190; the filter array is
commented out at Myfunctions.h:69 and its parameters unused at
viscoelas_3D.cpp:111-112), so nothing of the reference pins this mode; it is
pinned here by:
  * the factor itself against numpy's sinc and scipy's Tukey window;
  * the run: F = 1 reproduces the reference scheme bit for bit, F = 0 freezes
    every stress (so F reaches all five operator terms), the factorised
    restatement (A2 F as a 3-D operator) equals the literal oracle, the
    correction keeps a homogeneous medium stable at c dt / dx = 1, and it
    removes the time dispersion that grows with distance in the plain scheme.
"""
import numpy as np
import pytest
from scipy.signal.windows import tukey

import oracle as O
from cases import rel_l2, small_case

PI_T = 3.1415926  # the reference's truncated pi (cut-off wavenumber 2 pi fcut / vmax)


def test_sinc_squared_factor_matches_numpy():
    rng = np.random.default_rng(3)
    k = np.concatenate([[0.0], rng.uniform(0, 0.6, 2000)])
    for vref, dt in ((3000.0, 1e-3), (5000.0, 8e-4), (1500.0, 3.3e-3)):
        vh = 0.5 * float(np.float32(vref)) * float(np.float32(dt))
        want = np.sinc(vh * k / np.pi) ** 2
        got = O.kspace_factor(k, dt, vref=vref)
        assert got[0] == 1.0
        assert np.abs(got - want).max() < 1e-14
        got32 = O.kspace_factor(k, dt, vref=vref, prec="f32")
        assert np.abs(got32 - want).max() < 1e-6


@pytest.mark.parametrize("taper", [0.2, 0.5, 1.0])
def test_tukey_filter_matches_scipy(taper):
    """The radial taper on |k| in [0, kc] is the right half of scipy's Tukey
    window with alpha = taper_ratio; zero beyond kc."""
    fcut, vmax = 100.0, 5000.0
    kc = 2 * PI_T * float(np.float32(fcut)) / float(np.float32(vmax))
    M = 2 * 400 + 1
    w = tukey(M, alpha=float(np.float32(taper)))[400:]  # the ABI passes taper_ratio as FP32
    k = np.arange(401) / 400.0 * kc
    got = O.kspace_factor(k, 1e-3, fcut=fcut, vmax=vmax, taper=taper)
    assert np.abs(got - w).max() < 1e-12
    beyond = O.kspace_factor(kc * np.array([1.0001, 1.5, 3.0]), 1e-3, fcut=fcut, vmax=vmax, taper=taper)
    assert np.all(beyond == 0)
    both = O.kspace_factor(k, 1e-3, vref=4000.0, fcut=fcut, vmax=vmax, taper=taper)
    assert np.abs(both - w * O.kspace_factor(k, 1e-3, vref=4000.0)).max() < 1e-14


def test_hard_cut_and_bad_parameters():
    kc = 2 * PI_T * 50.0 / 2500.0
    k = np.array([0.0, 0.5 * kc, 0.999 * kc, 1.001 * kc, 2 * kc])
    assert np.array_equal(O.kspace_factor(k, 1e-3, fcut=50.0, vmax=2500.0, taper=0.0), [1, 1, 1, 0, 0])
    for bad in (dict(vref=-1.0), dict(fcut=10.0, vmax=0.0), dict(fcut=10.0, vmax=3000.0, taper=1.5)):
        with pytest.raises(ValueError):
            O.kspace_factor(k, 1e-3, **bad)


def test_unit_factor_is_the_reference_scheme_bitwise():
    """A cut-off beyond every |k| of the grid and no sinc: F = 1 exactly, and
    the run is the reference scheme bit for bit (FP32 and FP64)."""
    shot, model = small_case(16, 20, 24, 3, 30, O.SRC_EXPLOSIVE, True)
    for prec in ("f32", "f64"):
        a = O.run(shot, model, prec)["records"]
        b = O.run(shot, model, prec, kspace=dict(fcut=1e6, vmax=3000.0, taper=0.2))["records"]
        assert np.array_equal(a, b)


def test_zero_factor_freezes_every_stress():
    """A cut-off below the smallest nonzero |k|: F = 0 on every operator term
    (A1, A2, A3 on v and on dv), so the stresses and C-PML memory never move and
    only the source cell's velocity changes, by exactly the sum of the wavelet
    (lossy heterogeneous VTI model)."""
    shot, model = small_case(16, 20, 24, 3, 30, O.SRC_FZ, True)
    st = O.run(shot, model, "f64", state=True, kspace=dict(fcut=1e-3, vmax=3000.0, taper=0.0))["state"]
    assert not st[3:].any() and not st[0].any() and not st[1].any()
    src = (shot["sx"], shot["sy"], shot["sz"])
    assert np.count_nonzero(st[2]) == 1
    assert st[2][src] == shot["rik"].astype(np.float64).sum()


@pytest.mark.parametrize("case", [(10, 8, 12, 2, 10, O.SRC_EXPLOSIVE, True), (12, 10, 8, 3, 10, O.SRC_FZ, False)])
def test_factorised_kspace_scheme_equals_literal_oracle(case):
    """The engine's algebra in k-space mode (A2 F applied to v and dv as 3-D
    operators, F on A1 and A3) against the literal 45-transform oracle, FP64."""
    import factorised

    shot, model = small_case(*case)
    ks = dict(vref=3500.0, fcut=80.0, vmax=3000.0, taper=0.3)
    ref = O.run(shot, model, "f64", state=True, kspace=ks)
    rec, st = factorised.run(shot, model, case[6], kspace=ks)
    assert rel_l2(st, ref["state"]) < 1e-12
    assert rel_l2(rec, ref["records"]) < 1e-10


def _homogeneous(n, pml, nt, dt, f0, q=None, rx=None):
    shape = (n, n, n)
    vp = 3000.0
    g = lambda qq: 0.0 if q is None else np.arctan(1.0 / qq) / np.pi
    model = dict(vp=np.full(shape, vp, np.float32), vs=np.full(shape, vp / 1.73, np.float32),
                 rho=np.full(shape, 2000.0, np.float32), gp=np.full(shape, g(q), np.float32),
                 gs=np.full(shape, g(q / 2 if q else None), np.float32))
    c = n // 2
    rx = np.arange(c + 1, n - pml) if rx is None else np.asarray(rx)
    shot = dict(nx=n, ny=n, nz=n, pml=pml, nt=nt, dx=10.0, dy=10.0, dz=10.0, dt=dt, f0=25.0,
                w0=float(np.float32(2 * PI_T * 25.0)), src_kind=O.SRC_EXPLOSIVE, sx=c, sy=c, sz=c,
                rx=rx, ry=np.full(len(rx), c), rz=np.full(len(rx), c))
    # Ricker rate, injected as rate * dt at the step's effective time (it + 1/2) dt
    t = (np.arange(nt) + 0.5) * dt
    a = (np.pi * f0 * (t - 1.2 / f0)) ** 2
    shot["rik"] = (1 - 2 * a) * np.exp(-a) * dt / 1e-3
    return shot, model


@pytest.mark.parametrize("q", [None, 30.0])
def test_correction_is_stable_beyond_the_cfl_limit(q):
    """vp dt / dx = 1 (the plain scheme's limit is 2 / (pi sqrt 3) = 0.37):
    without the correction the run grows without bound, with vref a little
    above vp (Q dispersion raises the speed at high |k|) it stays bounded."""
    shot, model = _homogeneous(32, 4, 60, 10.0 / 3000.0, 25.0, q=q)
    shot["rik"] = shot["rik"].astype(np.float32)
    plain = O.run(shot, model, "f64")["records"]
    corr = O.run(shot, model, "f64", kspace=dict(vref=1.06 * 3000.0))["records"]
    assert np.isfinite(corr).all()
    peak = np.abs(corr).max()
    assert np.abs(corr[:, 40:]).max() <= peak
    assert not (np.abs(plain).max() < 1e6 * peak)


def test_correction_removes_time_dispersion():
    """Explosive source in a lossless homogeneous medium, 1.2 ms steps
    (vp dt / dx = 0.36) against 0.4 ms steps over the same physical times.
    The plain scheme's P-wave runs fast by ~(w dt)^2 / 24, so its coarse/fine
    misfit grows with distance; with vref = vp the leapfrog is exact in phase,
    and the misfit left (a near-field gain of the source's impulse train,
    compensated below to leading order) shrinks with distance.  sinc instead
    of sinc^2, or |k| dt instead of |k| dt / 2, fails the bound."""
    n, pml, f0, dtc, R = 48, 6, 40.0, 1.2e-3, 3
    ntc = 50
    rx = np.arange(30, 35)  # 6..10 cells from the source
    out = {}
    for name, ks in (("plain", None), ("corrected", dict(vref=3000.0))):
        recs = []
        for dt, nt in ((dtc, ntc), (dtc / R, ntc * R)):
            shot, model = _homogeneous(n, pml, nt, dt, f0, rx=rx)
            # per-mode impulse-train gain of the leapfrog, taken out of the
            # wavelet: w dt / sin(w dt) (corrected), 1 / cos(w dt / 2) (plain)
            L = 8 * nt
            spec = np.fft.rfft(shot["rik"], L)
            x = 2 * np.pi * np.fft.rfftfreq(L, dt) * dt
            h = np.sinc(x / np.pi) if ks else np.cos(x / 2)
            shot["rik"] = np.fft.irfft(spec * h, L)[:nt].astype(np.float32)
            recs.append(O.run(shot, model, "f64", kspace=ks)["records"][0])
        coarse, fine = recs[0], recs[1][1::R][:ntc]  # common times (n + 1/2) dtc
        out[name] = [rel_l2(coarse[:, r], fine[:, r]) for r in range(len(rx))]
    print("coarse/fine misfit per receiver:", out)
    assert max(out["corrected"]) < 0.013
    assert out["plain"][-1] > 4 * out["corrected"][-1]
    assert out["plain"][-1] > out["plain"][0] and out["corrected"][-1] < out["corrected"][0]
