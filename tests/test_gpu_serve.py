"""GPU tests of the continuous-batching server (qvti_server_*).

Serving must not change results: every request's seismograms are bit-identical
whether it runs alone or mixed with requests of other grid sizes, durations
and media (isotropic/VTI), and identical to the plain single-shot path
(qvti_forward), which the parity tests compare with the oracle.  Admission by
HBM budget, queueing when the budget is full, rejection of requests larger
than the budget, and cancellation are checked here too.
"""
import numpy as np
import pytest

import oracle as O
import qvti
from cases import rel_l2, small_case

pytestmark = pytest.mark.gpu

# nx, ny, nz, pml, nt, source kind, vti, seed
REQS = [
    (16, 20, 24, 3, 40, O.SRC_FX, False, 1),
    (24, 20, 16, 4, 30, O.SRC_FZ, True, 2),
    (20, 16, 32, 3, 55, O.SRC_EXPLOSIVE, False, 3),
    (16, 20, 24, 3, 25, O.SRC_FZ, True, 4),
    (20, 16, 32, 3, 35, O.SRC_FY, False, 5),   # same grid as the third: reuses a pooled context
    (32, 24, 20, 5, 45, O.SRC_EXPLOSIVE, True, 6),
]


def build(req):
    nx, ny, nz, pml, nt, kind, vti, seed = req
    shot, model = small_case(nx, ny, nz, pml, nt, kind, vti, seed=seed)
    grid = qvti.Grid(nx, ny, nz, pml, shot["dx"], shot["dy"], shot["dz"], shot["dt"], shot["f0"], shot["w0"],
                     int(vti))
    sh = qvti.make_shot(kind, (shot["sx"], shot["sy"], shot["sz"]), shot["rx"], shot["ry"], shot["rz"])
    return grid, shot, model, sh


def plain(req):
    grid, shot, model, sh = build(req)
    with qvti.Engine(grid.nx, grid.ny, grid.nz, grid.pml, grid.dx, grid.dy, grid.dz, grid.dt, grid.f0, grid.w0,
                     vti=bool(grid.vti)) as e:
        e.set_model_dict(model)
        return e.forward(shot["nt"], shot["rik"], sh)


def serve(reqs, budget, max_active, slice_steps):
    out, infos = [], []
    with qvti.Server(budget, max_active=max_active, steps_per_slice=slice_steps) as s:
        ids = []
        for r in reqs:
            grid, shot, model, sh = build(r)
            ids.append(s.submit(grid, shot["nt"], shot["rik"], model, sh))
        for i in ids:
            assert s.wait(i, 300.0) == 0, qvti.lib().qvti_last_error()
            rc, inf, rec = s.result(i)
            assert rc == 0
            out.append(rec)
            infos.append((inf.t_submit_s, inf.t_admit_s, inf.t_done_s))
    return out, infos


def estimate(req):
    grid = build(req)[0]
    return qvti.lib().qvti_estimate_bytes(grid, 1)


def test_each_result_is_the_same_alone_and_in_the_mix():
    big = 1 << 40
    alone = [serve([r], big, 1, 9)[0][0] for r in REQS]
    mixed, infos = serve(REQS, big, 4, 7)
    for r, a, m in zip(REQS, alone, mixed):
        assert np.array_equal(a, m), r
        assert np.array_equal(a, plain(r)), r
    # the mix really overlapped: some request was admitted before another finished
    assert any(infos[j][1] < infos[i][2] for i in range(len(infos)) for j in range(len(infos)) if i != j)


def test_served_records_match_the_oracle():
    r = REQS[1]
    grid, shot, model, sh = build(r)
    ref = O.run(shot, model, "f64")
    rec, _ = serve([REQS[0], r, REQS[2]], 1 << 40, 3, 5)
    la = shot["line_a"]
    assert rel_l2(rec[1][:, :, la:], ref["records"][:, :, la:]) < 1e-5


def test_served_request_on_a_serving_grid_matches_the_oracle():
    """A request on the serving stream's smallest grid (96^3, isotropic, Q 20
    to 200 drawn as in the stream, dx 12.5 m, dt 0.8 ms), served next to a
    VTI request: its first 12 steps against the FP32 / FP64 oracle."""
    n, pml, nt = 96, 20, 12
    m = qvti.make_model(2, 21, (n, n, n), [2000, 4000, 0.0, 20, 200, 15, 0.0])
    model = {k: m[k] for k in ("vp", "gp", "vs", "gs", "rho")}
    f0, dt, d = 20.0, 8e-4, 12.5
    w0 = float(np.float32(2 * 3.1415926 * f0))
    rx = np.arange(30, 66, 3)
    shot = dict(nx=n, ny=n, nz=n, pml=pml, nt=nt, dx=d, dy=d, dz=d, dt=dt, f0=f0, w0=w0, src_kind=O.SRC_FZ,
                sx=48, sy=48, sz=pml + 1, rx=rx, ry=np.full(len(rx), 48), rz=np.full(len(rx), pml + 1),
                rik=qvti.ricker(nt, dt, f0))
    grid = qvti.Grid(n, n, n, pml, d, d, d, dt, f0, w0, 0)
    sh = qvti.make_shot(O.SRC_FZ, (48, 48, pml + 1), shot["rx"], shot["ry"], shot["rz"])
    with qvti.Server(1 << 40, max_active=2, steps_per_slice=5) as s:
        g2, shot2, model2, sh2 = build(REQS[1])
        other = s.submit(g2, shot2["nt"], shot2["rik"], model2, sh2)
        i = s.submit(grid, nt, shot["rik"], model, sh)
        assert s.wait(i, 300.0) == 0 and s.wait(other, 300.0) == 0
        rc, _, rec = s.result(i)
        assert rc == 0
    r32 = O.run(shot, model, "f32")["records"]
    r64 = O.run(shot, model, "f64")["records"]
    e32, e64 = rel_l2(rec, r32), rel_l2(rec, r64)
    print("served 96^3 request, 12 steps: vs FP32 oracle %.2e, vs FP64 oracle %.2e" % (e32, e64))
    assert e32 < 1e-5 and e64 < 1e-5


def test_budget_queues_and_rejects():
    # room for exactly one context at a time: requests of different grids
    # are admitted one after another, evicting the pooled context
    reqs = [REQS[0], REQS[1], REQS[5]]
    budget = int(max(estimate(r) for r in reqs) * 1.05)
    out, infos = serve(reqs, budget, 4, 11)
    for r, rec in zip(reqs, out):
        assert np.array_equal(rec, plain(r)), r
    order = sorted(infos, key=lambda t: t[1])
    for a, b in zip(order, order[1:]):
        assert b[1] >= a[2] - 1e-6  # admitted only after the previous one finished
    # a request larger than the whole budget is rejected, the rest still run
    with qvti.Server(budget, max_active=2) as s:
        grid, shot, model, sh = build((64, 64, 64, 4, 10, O.SRC_FX, True, 7))
        bad = s.submit(grid, shot["nt"], shot["rik"], model, sh)
        grid, shot, model, sh = build(REQS[0])
        good = s.submit(grid, shot["nt"], shot["rik"], model, sh)
        assert s.wait(bad, 60.0) == qvti.E_BUDGET
        assert s.wait(good, 60.0) == 0


def test_cancel():
    with qvti.Server(1 << 40, max_active=2, steps_per_slice=5) as s:
        grid, shot, model, sh = build((32, 32, 32, 4, 4000, O.SRC_FX, False, 8))
        rik = np.resize(shot["rik"], 4000) * 0  # a silent source: stable for any length
        long_id = s.submit(grid, 4000, rik, model, sh)
        grid, shot, model, sh = build(REQS[0])
        short_id = s.submit(grid, shot["nt"], shot["rik"], model, sh)
        s.cancel(long_id)
        assert s.wait(long_id, 120.0) == qvti.E_CANCELLED
        assert s.wait(short_id, 120.0) == 0
        rc, _, rec = s.result(short_id)
        assert rc == 0 and np.array_equal(rec, plain(REQS[0]))


def _serve_inputs(k, nt=None):
    """Grid, model, shot and wavelet of serving-grid request k
    (tests/fullsize_cases.py), optionally with another step count."""
    import fullsize_cases as FC

    g, m, shot = FC.serve_case(k)
    nt = nt or shot["nt"]
    rik = shot["rik"] if nt == shot["nt"] else qvti.ricker(nt, g["dt"], g["f0"])
    grid = qvti.Grid(g["nx"], g["ny"], g["nz"], g["pml"], g["dx"], g["dy"], g["dz"], g["dt"], g["f0"], g["w0"],
                     int(g["vti"]))
    return grid, m, FC.make_shot(shot), nt, rik


def test_every_serving_grid_matches_its_fixture_in_a_mix():
    """One request on each serving grid (96^3 isotropic, 128^3 VTI,
    160x160x128 isotropic, 200x200x100 VTI; 150 steps of 0.8 ms; receivers
    off the source cell), all four in flight together: each gather within
    1e-5 of the FP32 and FP64 oracle fixtures (tests/golden/serve_grids.npz)."""
    import os

    import fullsize_cases as FC

    gold = np.load(os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "serve_grids.npz"))
    with qvti.Server(1 << 40, max_active=4, steps_per_slice=25) as s:
        ids = []
        for k in range(len(FC.SERVE_CASES)):
            grid, m, sh, nt, rik = _serve_inputs(k)
            assert FC.model_checksum(m) == gold["model_sum_%d" % k]
            ids.append(s.submit(grid, nt, rik, m, sh))
        for k, i in enumerate(ids):
            assert s.wait(i, 300.0) == 0, qvti.lib().qvti_last_error()
            rc, _, rec = s.result(i)
            assert rc == 0
            e32 = rel_l2(rec, gold["rec_f32_%d" % k])
            e64 = rel_l2(rec, gold["rec_f64_%d" % k])
            print("served %s %s: vs FP32 oracle %.2e, vs FP64 oracle %.2e" % (
                "x".join(map(str, FC.SERVE_CASES[k][0])), "VTI" if FC.SERVE_CASES[k][1] else "iso", e32, e64),
                flush=True)
            assert e32 < 1e-5 and e64 < 1e-5, k


def test_long_request_is_the_same_alone_and_in_the_mix():
    """A 3000-step request on the 96^3 serving grid (the serving stream's
    longest duration) is bit-identical served alone and served while requests
    on the 128^3 and 200x200x100 grids come and go."""
    long_req = _serve_inputs(0, nt=3000)
    with qvti.Server(1 << 40, max_active=1, steps_per_slice=50) as s:
        grid, m, sh, nt, rik = long_req
        i = s.submit(grid, nt, rik, m, sh)
        assert s.wait(i, 300.0) == 0
        alone = s.result(i)[2]
    with qvti.Server(1 << 40, max_active=3, steps_per_slice=50) as s:
        grid, m, sh, nt, rik = long_req
        i = s.submit(grid, nt, rik, m, sh)
        others = []
        for k in (1, 3):
            g2, m2, sh2, nt2, rik2 = _serve_inputs(k)
            others.append(s.submit(g2, nt2, rik2, m2, sh2))
        for o in others:
            assert s.wait(o, 300.0) == 0
        assert s.wait(i, 300.0) == 0
        infos = [s.info(o)[1] for o in others] + [s.info(i)[1]]
        mixed = s.result(i)[2]
    assert np.isfinite(alone).all()
    assert np.array_equal(alone, mixed)
    # the other requests were in flight together with the long one
    long_inf = infos[-1]
    assert all(inf.t_admit_s < long_inf.t_done_s and long_inf.t_admit_s < inf.t_done_s for inf in infos[:-1])
