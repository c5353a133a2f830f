"""Full-size parity cases shared by the fixture generator
(tests/golden/make_fullsize_golden.py, run once in the build container) and
the GPU tests that check the engine against those fixtures.

Every case regenerates its model from a seed through qvti_make_model (the
library's host-side generator, identical on any machine) and records at
receivers that exclude the source cell, over a window long enough for the
wavefront to reach receivers 10-40 cells away and to enter the C-PML.

The time step is 0.5 ms (0.8 ms for the serving grids): the configurations'
specified 1 ms exceeds the staggered pseudospectral stability limit
c dt / dx <= 2 / (pi sqrt(3)) for their fastest horizontal qP speeds, and the
scheme (the reference's too) grows without bound there.
"""
import numpy as np

import qvti

PI = 3.1415926

# name: (nx, ny, nz, nt)
FULLSIZE = {
    "vti256": (256, 256, 256, 200),
    "vti400": (400, 400, 200, 160),
}
PML, DX, DT, F0 = 20, 12.5, 5e-4, 20.0
# seeded smooth random VTI model of the 256^3 / 400x400x200 / 160x160x128
# configurations: vp in [2000, 5000], Qp in [20, 150], eps in [0.05, 0.25]
# falling with vp, delta = eps / 2, Gaussian correlation sigma 15 cells
SMOOTH_PARAMS = [2000, 5000, 0.25, 20, 150, 15, 0.05]
MODEL_SEED = 7


def wavelet(nt, dt=DT, f0=F0, delay=0.75):
    """Ricker of peak frequency f0 delayed by delay/f0 seconds, FP64-evaluated
    then rounded (the reference's getrik with delay 1/f0,
    viscoelas_3D.cpp:494-502, moved earlier so that the pulse is emitted within
    a short window)."""
    t = np.arange(nt, dtype=np.float64) * dt
    a = (PI * f0 * (t - delay / f0)) ** 2
    return ((1.0 - 2.0 * a) * np.exp(-a)).astype(np.float32)


def grid_dict(nx, ny, nz, dt=DT, vti=True, pml=PML, dx=DX, f0=F0):
    return dict(nx=nx, ny=ny, nz=nz, pml=pml, dx=dx, dy=dx, dz=dx, dt=dt, f0=f0,
                w0=float(np.float32(2 * PI * f0)), vti=vti)


def fullsize_receivers(nx, ny, nz, sx, sy, sz, pml=PML):
    """Line A along x just below the top C-PML; line B vertical, 12 cells off
    the source; line C along y at the source depth, 10-40 cells away; three
    points inside the x and top C-PML.  None is the source cell."""
    pts = []
    for k in range(-10, 11):
        pts.append((sx + 4 * k, sy, pml + 2))
    for z in range(2, min(nz - 3, sz + 61), 6):
        pts.append((sx + 12, sy, z))
    for d in range(10, 41, 5):
        pts.append((sx, sy + d, sz))
    pts += [(4, sy, sz), (10, sy, sz), (sx, sy, 8)]
    r = np.array(pts, np.int32)
    assert not np.any((r[:, 0] == sx) & (r[:, 1] == sy) & (r[:, 2] == sz))
    assert r.min() >= 0 and r[:, 0].max() < nx and r[:, 1].max() < ny and r[:, 2].max() < nz
    return r[:, 0].copy(), r[:, 1].copy(), r[:, 2].copy()


def fullsize_case(name):
    """(grid, model, shot) of a full-size single-shot case: vz force at
    (nx/2, ny/2, pml + 6)."""
    nx, ny, nz, nt = FULLSIZE[name]
    g = grid_dict(nx, ny, nz)
    m = qvti.make_model(2, MODEL_SEED, (nx, ny, nz), SMOOTH_PARAMS, vti=True)
    src = (nx // 2, ny // 2, PML + 6)
    rx, ry, rz = fullsize_receivers(nx, ny, nz, *src)
    shot = dict(kind=qvti.SRC_FZ, src=src, rx=rx, ry=ry, rz=rz, nt=nt, rik=wavelet(nt))
    return g, m, shot


# ---- batched configuration: 32 shots on one 160x160x128 model
BATCH_SHAPE, BATCH_SHOTS, BATCH_NT, BATCH_SEED = (160, 160, 128), 32, 200, 11


def batch_case():
    """(grid, model, shots) of the batched configuration: sources at seeded
    uniform (x, y) inside the C-PML frame at z = pml + 1, every shot with a
    receiver line along x through its source's y at z = pml + 4, from 30
    cells before to 30 cells after the source (clipped to the grid; the
    source cells are not on it)."""
    nx, ny, nz = BATCH_SHAPE
    g = grid_dict(nx, ny, nz)
    m = qvti.make_model(2, MODEL_SEED, (nx, ny, nz), SMOOTH_PARAMS, vti=True)
    rng = np.random.default_rng(BATCH_SEED)
    xs = rng.integers(PML, nx - PML, BATCH_SHOTS)
    ys = rng.integers(PML, ny - PML, BATCH_SHOTS)
    shots = []
    for x, y in zip(xs, ys):
        rx = np.arange(max(0, int(x) - 30), min(nx, int(x) + 31), 3, dtype=np.int32)
        shots.append(dict(kind=qvti.SRC_FZ, src=(int(x), int(y), PML + 1), rx=rx,
                          ry=np.full(len(rx), int(y), np.int32), rz=np.full(len(rx), PML + 4, np.int32)))
    return g, m, shots


# ---- serving grids (one request each)
SERVE_CASES = [((96, 96, 96), False, 101), ((128, 128, 128), True, 102), ((160, 160, 128), False, 103),
               ((200, 200, 100), True, 104)]
SERVE_NT, SERVE_DT = 150, 8e-4
# vp in [2000, 4000], Qp in [20, 200], eps in [0.05, 0.2], sigma 8 cells
SERVE_PARAMS = [2000, 4000, 0.2, 20, 200, 8, 0.05]


def serve_case(k):
    """(grid, model, shot) of serving request k: vz force at (nx/2, ny/2,
    pml + 1), receivers along x at y = ny/2 + 3, z = pml + 4."""
    (nx, ny, nz), vti, seed = SERVE_CASES[k]
    g = grid_dict(nx, ny, nz, dt=SERVE_DT, vti=vti)
    m = qvti.make_model(2, seed, (nx, ny, nz), SERVE_PARAMS, vti=vti)
    if not vti:
        m.pop("eps", None)
        m.pop("delta", None)
    rx = np.arange(0, nx, 2, dtype=np.int32)
    shot = dict(kind=qvti.SRC_FZ, src=(nx // 2, ny // 2, PML + 1), rx=rx, ry=np.full(len(rx), ny // 2 + 3, np.int32),
                rz=np.full(len(rx), PML + 4, np.int32), nt=SERVE_NT, rik=qvti.ricker(SERVE_NT, SERVE_DT, F0))
    return g, m, shot


def oracle_inputs(g, m, shot, nt, rik):
    """The oracle's (shot, model) dicts for one case."""
    s = dict(nx=g["nx"], ny=g["ny"], nz=g["nz"], pml=g["pml"], nt=nt, dx=g["dx"], dy=g["dy"], dz=g["dz"],
             dt=g["dt"], f0=g["f0"], w0=g["w0"], src_kind=shot["kind"], sx=shot["src"][0], sy=shot["src"][1],
             sz=shot["src"][2], rx=shot["rx"], ry=shot["ry"], rz=shot["rz"], rik=rik)
    model = {k: m[k] for k in ("vp", "gp", "vs", "gs", "rho")}
    if g["vti"]:
        model["eps"], model["delta"] = m["eps"], m["delta"]
    return s, model


def model_checksum(m):
    return np.float64(sum(m[k].astype(np.float64).sum() for k in ("vp", "gp", "vs", "gs", "rho")))


def engine_for(g, max_shots=1):
    return qvti.Engine(g["nx"], g["ny"], g["nz"], g["pml"], g["dx"], g["dy"], g["dz"], g["dt"], g["f0"], g["w0"],
                       vti=g["vti"], max_shots=max_shots)


def make_shot(shot):
    return qvti.make_shot(shot["kind"], shot["src"], shot["rx"], shot["ry"], shot["rz"])
