"""Shared synthetic test cases (small grids the oracle finishes in seconds)."""
import numpy as np

PI = 3.1415926


def small_case(nx, ny, nz, pml, nt, kind, vti, seed=1, dt=5e-4, dx=10.0, dy=12.0, dz=9.0, f0=30.0):
    """Random heterogeneous model + shot dict in the oracle's format."""
    rng = np.random.default_rng(seed)
    shape = (nx, ny, nz)
    sx, sy, sz = nx // 2, ny // 2 - 1, nz // 2 + 1
    # line A: along x just below the top C-PML (far field, small amplitudes);
    # line B: along y through the source column (near field)
    rxa = np.arange(1, nx - 1, 2)
    ryb = np.arange(0, ny)
    rx = np.concatenate([rxa, np.full(len(ryb), sx + 1)])
    ry = np.concatenate([np.full(len(rxa), ny // 2 + 1), ryb])
    rz = np.concatenate([np.full(len(rxa), pml + 1), np.full(len(ryb), sz - 2)])
    shot = dict(nx=nx, ny=ny, nz=nz, pml=pml, nt=nt, dx=dx, dy=dy, dz=dz, dt=dt, f0=f0,
                w0=float(np.float32(2 * PI * f0)), src_kind=kind, sx=sx, sy=sy, sz=sz,
                rx=rx, ry=ry, rz=rz, line_a=len(rxa))
    t = np.arange(nt)
    shot["rik"] = (np.sin(t * 0.3) * np.exp(-0.01 * t)).astype(np.float32)
    vp = (2500 + 500 * rng.random(shape)).astype(np.float32)
    model = dict(vp=vp, vs=(vp / 1.8).astype(np.float32),
                 rho=(2000 + 300 * rng.random(shape)).astype(np.float32),
                 gp=(0.002 + 0.01 * rng.random(shape)).astype(np.float32),
                 gs=(0.004 + 0.01 * rng.random(shape)).astype(np.float32))
    if vti:
        model["eps"] = (0.2 * rng.random(shape)).astype(np.float32)
        model["delta"] = (model["eps"] * rng.random(shape)).astype(np.float32)
    return shot, model


def rel_l2(a, b):
    a = np.asarray(a, np.float64)
    b = np.asarray(b, np.float64)
    return float(np.linalg.norm(a - b) / max(np.linalg.norm(b), 1e-300))
