"""GPU tests of the opt-in k-space mode (qvti_set_kspace): parity with the
oracle's k-space mode, the F = 0 invariant, bit-identity across tile widths
and batching, the stabilisation of the layered VTI 128^3 configuration at its
specified 1 ms step (beyond the plain scheme's stability limit), and argument
checks."""
import numpy as np
import pytest

import oracle as O
import qvti
from cases import small_case
from test_gpu_engine import check_parity, engine_for, shot_of

pytestmark = pytest.mark.gpu

KS = dict(vref=3600.0, fcut=60.0, vmax=3000.0, taper=0.3)


def ks_args(ks):
    return ks.get("vref", 0.0), ks.get("fcut", 0.0), ks.get("vmax", 0.0), ks.get("taper", 0.0)


@pytest.mark.parametrize("case", [(24, 20, 32, 4, 30, O.SRC_EXPLOSIVE, True), (32, 24, 20, 5, 30, O.SRC_FX, False),
                                  (20, 16, 32, 3, 30, O.SRC_FZ, True)])
def test_kspace_parity_vs_oracle(case):
    """sinc^2 correction and Tukey filter both on (cut-off inside the grid's
    wavenumber range): state and records against the FP32 / FP64 oracle in
    the same mode, to the engine's usual bounds."""
    nx, ny, nz, pml, nt, kind, vti = case
    shot, model = small_case(nx, ny, nz, pml, nt, kind, vti)
    ref32 = O.run(shot, model, "f32", state=True, kspace=KS)
    ref64 = O.run(shot, model, "f64", state=True, kspace=KS)
    plain = O.run(shot, model, "f64")["records"]
    with engine_for(shot, model) as e:
        e.set_kspace(*ks_args(KS))
        e.begin(nt, shot["rik"], [shot_of(shot)])
        e.step(nt)
        rec = e.finish()[0]
        st = e.state()
    assert not np.allclose(plain, ref64["records"], rtol=1e-3)  # the mode changes the result
    check_parity(case, rec, st, ref32, ref64, shot["line_a"])


def test_zero_factor_freezes_every_stress():
    """F = 0 for every nonzero |k|: the stresses stay at zero and only the
    source cell's vz moves, by the running sum of the wavelet.  The oracle
    gets exact zeros (F multiplies each term's spectrum); the engine applies
    A2 F as a real field - here the constant DC part of v - and then a 1-D
    derivative transform, whose FP32 rounding leaves residues ~1e-7 of that
    constant in the non-DC bins, so the bound is relative to the scale a
    plain run reaches (sigma ~ rho c v)."""
    shot, model = small_case(16, 20, 24, 3, 30, O.SRC_FZ, True)
    with engine_for(shot, model) as e:
        e.set_kspace(0.0, 1e-3, 3000.0, 0.0)
        e.begin(shot["nt"], shot["rik"], [shot_of(shot)])
        e.step(shot["nt"])
        e.finish()
        st = e.state()
        e.set_kspace(0.0, 0.0, 0.0, 0.0)
        e.begin(shot["nt"], shot["rik"], [shot_of(shot)])
        e.step(shot["nt"])
        e.finish()
        plain = e.state()
    src = (shot["sx"], shot["sy"], shot["sz"])
    vsum = float(shot["rik"].astype(np.float64).sum())
    assert abs(st[2][src] - vsum) <= 1e-6 * abs(vsum)
    v = st[:3].copy()
    v[2][src] = 0
    print("max |v| off the source %.2e, max |sigma| %.2e (plain run %.2e)"
          % (np.abs(v).max(), np.abs(st[3:]).max(), np.abs(plain[3:]).max()))
    assert np.abs(v).max() < 1e-9 * abs(vsum)
    assert np.abs(st[3:]).max() < 1e-9 * np.abs(plain[3:]).max()


@pytest.mark.parametrize("mode", [0, 2, 20, 1])
def test_kspace_tile_widths_and_batching_bitwise(mode):
    """Every tile choice (the y-inverse pass has 8 fields in this mode) and a
    2-shot batch give the same bits as the narrow single-shot run."""
    shot, model = small_case(64, 96, 40, 6, 10, O.SRC_EXPLOSIVE, True)
    with engine_for(shot, model, max_shots=2) as e:
        e.set_kspace(*ks_args(KS))
        e.set_autotune(mode)
        batch = e.forward_batch(shot["nt"], shot["rik"], [shot_of(shot), shot_of(shot, sx=20)])
    with engine_for(shot, model) as e:
        e.set_autotune(0)
        e.set_kspace(*ks_args(KS))
        ref = e.forward(shot["nt"], shot["rik"], shot_of(shot))
        # switching the mode off again restores the plain scheme
        e.set_kspace(0.0, 0.0, 0.0, 0.0)
        plain = e.forward(shot["nt"], shot["rik"], shot_of(shot))
    with engine_for(shot, model) as e:
        e.set_autotune(0)
        plain0 = e.forward(shot["nt"], shot["rik"], shot_of(shot))
    assert np.array_equal(batch[0], ref)
    assert np.array_equal(plain, plain0)
    assert not np.array_equal(plain, ref)


def test_kspace_stabilises_layered_vti_at_1ms():
    """The layered VTI 128^3 configuration (vp up to 4500 m/s, eps up to 0.25,
    Qp down to 30) at its specified dt = 1 ms and full length (2000 steps):
    c dt / dx reaches 0.55 against the plain scheme's limit 0.37, and the
    blow-up probe stops the plain run; with vref = 1.25 x the fastest
    horizontal qP speed (Q dispersion adds ~12% at the grid's highest |k|)
    the run stays finite and bounded (|v| of a unit force source stays far
    below the probe's limit)."""
    n, pml, f0, dt, nt = 128, 20, 20.0, 1e-3, 2000
    m = qvti.make_model(1, 11, (n, n, n), [1800, 4500, 0.25, 30, 200], nlay=6, vti=True)
    vh = float((m["vp"].astype(np.float64) * np.sqrt(1 + 2 * m["eps"].astype(np.float64))).max())
    rik = qvti.ricker(nt, dt, f0)
    rx = np.arange(20, 108, 4)
    shot = qvti.make_shot(qvti.SRC_FZ, (64, 64, 10), rx, np.full(len(rx), 64), np.full(len(rx), pml))
    w0 = float(np.float32(2 * 3.1415926 * f0))
    with qvti.Engine(n, n, n, pml, 10.0, 10.0, 10.0, dt, f0, w0, vti=True) as e:
        e.set_model_dict(m)
        e.set_probe(10, 1e4)
        with pytest.raises(qvti.QvtiError) as err:
            e.forward(nt, rik, shot)
        assert err.value.code == qvti.E_UNSTABLE
        e.set_kspace(1.25 * vh, 0.0, 0.0, 0.0)
        rec = e.forward(nt, rik, shot)
    assert np.isfinite(rec).all()
    peak = np.abs(rec).max()
    assert 0 < peak < 1e2


def test_set_kspace_rejects_bad_arguments():
    shot, model = small_case(16, 16, 16, 3, 5, O.SRC_FX, False)
    with engine_for(shot, model) as e:
        for bad in ((-1.0, 0.0, 0.0, 0.0), (0.0, 10.0, 0.0, 0.2), (0.0, 10.0, 3000.0, 1.5), (float("nan"), 0, 0, 0)):
            with pytest.raises(qvti.QvtiError) as err:
                e.set_kspace(*bad)
            assert err.value.code == qvti.E_ARG


def test_kspace_full_length_shot_at_256():
    """The 256^3 smooth VTI configuration as specified: 3000 steps of 1 ms
    (c dt / dx up to 0.49 before Q dispersion).  With the k-space correction
    (vref = 1.25 x the fastest horizontal qP speed) the whole shot runs with
    the blow-up probe armed, and the gathers stay finite and bounded."""
    import bench

    g, m, sh = bench.make_workload("vti256")
    nt = 3000
    vh = float((m["vp"].astype(np.float64) * np.sqrt(1 + 2 * m["eps"].astype(np.float64))).max())
    rik = qvti.ricker(nt, g["dt"], g["f0"])
    with bench.engine(g) as e:
        e.set_model_dict(m)
        e.set_probe(50, 1e4)
        e.set_kspace(1.25 * vh)
        rec = e.forward(nt, rik, qvti.make_shot(sh["kind"], sh["src"], sh["rx"], sh["ry"], sh["rz"]))
    assert np.isfinite(rec).all()
    peak = np.abs(rec).max()
    print("256^3, 3000 steps, k-space mode: peak |v| at the receivers %.3e" % peak)
    assert 0 < peak < 1e2
