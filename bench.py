"""Benchmark of the MI355X Q-VTI engine.

Default: the plain single-shot path on the 256^3 VTI workload (total grid,
20-cell C-PML, dx = 12.5 m, dt = 1 ms, seeded smooth random model), W warm-up
steps then K timed steps.  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload vti256|vti400|vti128|iso64]
                  [--batched] [--served] [--no-cpu-baseline]

value = total cells x K / timed wall seconds / 1e9 (Gcell-steps/s).  The
timed region brackets K steps with a device synchronisation on both sides;
every kernel launch inside it is bracketed by HIP events on the engine's
stream, from which the dominant kernel's average duration and achieved
algorithmic GB/s are computed.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("q-vti_amd", "oracle", "tests"):
    sys.path.insert(0, os.path.join(ROOT, sub))

import qvti  # noqa: E402

PI = 3.1415926
HBM_PEAK_GBS = 8000.0

# name: (nx, ny, nz, pml, dx, dt, vti)
WORKLOADS = {
    "iso64": (64, 64, 64, 10, 10.0, 1e-3, False),
    "vti128": (128, 128, 128, 20, 10.0, 1e-3, True),
    "vti256": (256, 256, 256, 20, 12.5, 1e-3, True),
    "vti400": (400, 400, 200, 20, 12.5, 1e-3, True),
    "vti160": (160, 160, 128, 20, 12.5, 1e-3, True),
}


def b_alg(nx, ny, nz, pml, vti):
    """Algorithmic bytes per cell-step of the scheme (FP32): 9 forward and 54
    inverse 3-D real transforms at 8 B/cell each, 18 dt-terms reading the
    previous spectrum, the velocity and stress updates with C coefficient
    arrays, and the C-PML memory variables on 2*pml-thick face slabs."""
    C = 30 if vti else 15
    slab = sum(2.0 * pml / n for n in (nx, ny, nz))
    return 72 + 432 + 72 + 72 + (180 + 48 + 4 * C) + 64 * slab


def make_workload(name, seed=7):
    nx, ny, nz, pml, d, dt, vti = WORKLOADS[name]
    if vti:
        # smooth random vp in [2000, 5000], Q in [20, 150], eps in [0.05, 0.25] co-varying
        m = qvti.make_model(2, seed, (nx, ny, nz), [2000, 5000, 0.25, 20, 150, 15, 0.05], vti=True)
    else:
        m = qvti.make_model(0, seed, (nx, ny, nz), [3000, 1730, 2200, 100, 50])
    f0 = 25.0 if name == "iso64" else 20.0
    g = dict(nx=nx, ny=ny, nz=nz, pml=pml, dx=d, dy=d, dz=d, dt=dt, f0=f0, w0=float(np.float32(2 * PI * f0)), vti=vti)
    rx = np.arange(nx)
    src = (nx // 2, ny // 2, pml + 1)
    shot = dict(kind=qvti.SRC_FZ, src=src, rx=rx, ry=np.full(nx, ny // 2), rz=np.full(nx, pml + 1))
    return g, m, shot


def engine(g, max_shots=1):
    return qvti.Engine(g["nx"], g["ny"], g["nz"], g["pml"], g["dx"], g["dy"], g["dz"], g["dt"], g["f0"], g["w0"],
                       vti=g["vti"], max_shots=max_shots)


def run_plain(name, steps, warmup):
    g, m, sh = make_workload(name)
    nt = warmup + steps
    rik = qvti.ricker(nt, g["dt"], g["f0"])
    e = engine(g)
    t0 = time.time()
    e.set_model_dict(m)
    setup_s = time.time() - t0
    e.set_graphs(False)
    e.set_probe(0)
    e.begin(nt, rik, [qvti.make_shot(sh["kind"], sh["src"], sh["rx"], sh["ry"], sh["rz"])])
    e.step(warmup)
    e.sync()
    e.enable_kernel_timing(steps)
    e.sync()
    t0 = time.perf_counter()
    e.step(steps)
    e.sync()
    wall = time.perf_counter() - t0
    kt = e.kernel_times()
    e.close()
    return g, wall, kt, setup_s


def run_cpu_baseline(name, budget_s=25.0):
    """FP32 CPU oracle (literal restatement of the reference scheme, OpenMP on
    all host cores) timed on a bounded number of steps of the workload."""
    import oracle as O

    g, m, sh = make_workload(name)
    cores = os.cpu_count() or 1
    cores = min(cores, int(os.environ.get("OMP_NUM_THREADS", cores)))
    N = g["nx"] * g["ny"] * g["nz"]
    shot = dict(nx=g["nx"], ny=g["ny"], nz=g["nz"], pml=g["pml"], dx=g["dx"], dy=g["dy"], dz=g["dz"], dt=g["dt"],
                f0=g["f0"], w0=g["w0"], src_kind=sh["kind"], sx=sh["src"][0], sy=sh["src"][1], sz=sh["src"][2],
                rx=sh["rx"], ry=sh["ry"], rz=sh["rz"])
    model = {k: m[k] for k in ("vp", "gp", "vs", "gs", "rho")}
    if g["vti"]:
        model["eps"], model["delta"] = m["eps"], m["delta"]
    # one step first; then as many as fit in the budget
    nt = 1
    shot["nt"] = nt
    shot["rik"] = qvti.ricker(nt, g["dt"], g["f0"])
    t0 = time.perf_counter()
    O.run(shot, model, "f32", nthreads=cores)
    t1 = time.perf_counter() - t0
    per = t1  # includes one-off setup; refine with a 2-step run when cheap
    steps = 1
    if 3 * t1 < budget_s:
        shot["nt"] = 3
        shot["rik"] = qvti.ricker(3, g["dt"], g["f0"])
        t0 = time.perf_counter()
        O.run(shot, model, "f32", nthreads=cores)
        t3 = time.perf_counter() - t0
        per = (t3 - t1) / 2 if t3 > t1 else t3 / 3
        steps = 3
    return {"value": N / per / 1e9, "unit": "Gcell-steps/s", "cores": cores, "kind": "port",
            "sample": "%s FP32 oracle, %d-step run(s), per-step time from the difference of a 3-step and a "
                      "1-step run (setup excluded)" % (name, steps)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="vti256", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (shots shard across GPUs with no communication; "
              "see DESIGN.md); --gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    name = args.workload
    g, wall, kt, setup_s = run_plain(name, args.steps, args.warmup)
    N = g["nx"] * g["ny"] * g["nz"]
    gcs = N * args.steps / wall / 1e9
    balg = b_alg(g["nx"], g["ny"], g["nz"], g["pml"], g["vti"])
    # dominant kernel
    best = max((k for k in kt if kt[k]["launches"] > 0), key=lambda k: kt[k]["total_ms"])
    k = kt[best]
    avg_ms = k["total_ms"] / k["launches"]
    achieved = k["bytes_per_launch"] / (avg_ms * 1e-3) / 1e9
    per_kernel = {n: {"avg_us": round(1e3 * v["total_ms"] / v["launches"], 2),
                      "share": round(v["total_ms"] / sum(x["total_ms"] for x in kt.values()), 4),
                      "GBps": round(v["bytes_per_launch"] / (v["total_ms"] / v["launches"] * 1e-3) / 1e9, 1)}
                  for n, v in kt.items() if v["launches"] > 0}
    out = {
        "metric": "Gcell-steps/s at 256^3 VTI FP32 (%HBM roofline); batched shots/s; served p99",
        "value": round(gcs, 4),
        "unit": "Gcell-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(wall / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic (seeded smooth random VTI model, vz-force Ricker source)",
        "config": {"workload": "%s single shot: %dx%dx%d total grid, pml %d, dx %.1f m, dt %.0e s, %s FP32"
                               % (name, g["nx"], g["ny"], g["nz"], g["pml"], g["dx"], g["dt"],
                                  "VTI" if g["vti"] else "isotropic"),
                   "cells": N, "parallelism": "1 GPU, 1 shot"},
        "roofline": {"bound": "hbm", "kernel": best, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": None},
        "step_roofline": {"b_alg_bytes_per_cell_step": round(balg, 1),
                          "achieved_GBps": round(balg * gcs, 1),
                          "frac": round(balg * gcs / HBM_PEAK_GBS, 4)},
        "kernels": per_kernel,
        "setup_s": round(setup_s, 3),
    }
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = run_cpu_baseline(name)
        except Exception as ex:  # the GPU number stands on its own
            out["cpu_baseline"] = {"error": str(ex)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
