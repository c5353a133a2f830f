"""Benchmark of the MI355X Q-VTI engine.

Default: the plain single-shot path on the 256^3 VTI workload (total grid,
20-cell C-PML, dx = 12.5 m, dt = 1 ms, seeded smooth random model), W warm-up
steps then K timed steps, followed by the other legs of the metric - the
32-shot batched configuration, a seeded Poisson serving stream, a stream-copy
bandwidth probe - and the CPU baseline.  Prints ONE JSON line.

  python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload vti256|vti400|vti160|vti128|iso64]
                  [--batched] [--served NREQ] [--tiles MODE] [--kspace VREF] [--no-cpu-baseline]
                  [--headline-only] [--copy-probe]
  python bench.py --workload W --summarize-counters TRACE FETCH WRITE COPY_FETCH COPY_WRITE BENCH_JSON
  python bench.py --summarize-pmc DIR [DIR ...]

value = total cells x K / timed wall seconds / 1e9 (Gcell-steps/s) over K
steps of the production path (HIP graph per step), bracketed by device
synchronisations.  A second region of K steps launches the same kernels
directly with every launch bracketed by HIP events on the engine's stream;
the dominant kernel's average duration there gives the achieved algorithmic
GB/s of the roofline object.  roofline.traffic is the measured HBM bytes per
launch of that kernel from the committed rocprofv3 counter summary, used only
while the kernel sources still hash to the ones that summary was measured on.
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


def run_plain(name, steps, warmup, kspace_vref=0.0, tiles=1):
    """Two timed regions of `steps` steps each on one shot, after `warmup`
    steps: (1) the production path (one HIP graph per step, blow-up probe
    every 50 steps) gives the throughput; (2) the same kernels launched
    directly, each bracketed by HIP events on the engine's stream, gives the
    per-kernel average durations.  kspace_vref > 0 runs the opt-in k-space
    corrected scheme instead."""
    g, m, sh = make_workload(name)
    nt = warmup + 2 * steps
    rik = qvti.ricker(nt, g["dt"], g["f0"])
    e = engine(g)
    t0 = time.time()
    e.set_model_dict(m)
    if kspace_vref > 0:
        e.set_kspace(kspace_vref)
    if tiles != 1:
        e.set_autotune(tiles)
    setup_s = time.time() - t0
    t0 = time.perf_counter()
    e.begin(nt, rik, [qvti.make_shot(sh["kind"], sh["src"], sh["rx"], sh["ry"], sh["rz"])])
    begin_s = time.perf_counter() - t0
    tiles = {"pencils": e.tile_choice(), "split": e.tile_split()}
    blow_up = []

    def sync():
        # qvti_sync waits for the stream before it checks the blow-up probe,
        # so the timing stands when the probe fires: a workload whose dt is
        # beyond the scheme's stability limit still does the same work per step
        try:
            e.sync()
        except qvti.QvtiError as ex:
            if ex.code != qvti.E_UNSTABLE:
                raise
            if not blow_up:
                blow_up.append(e.steps_done())

    e.step(warmup)
    sync()
    t0 = time.perf_counter()
    e.step(steps)
    sync()
    wall = time.perf_counter() - t0
    e.enable_kernel_timing(steps)
    sync()
    t0 = time.perf_counter()
    e.step(steps)
    sync()
    wall_timed = time.perf_counter() - t0
    kt = e.kernel_times()
    e.close()
    return g, wall, wall_timed, kt, setup_s, begin_s, tiles, (blow_up[0] if blow_up else None)


KERNEL_SOURCES = ("qvti_engine.hip", "qvti_ctx.h", "qvti_kernels.hip", "qvti_launch.h", "qvti_lengths.h",
                  "qvti_layout.h", "qvti_pencils.h", "qvti_fft2l.h", "qvti_aux.h")


def kernel_source_hash():
    """sha256 (first 16 hex digits) of the engine's kernel and launch sources:
    a committed counter summary is only used while it matches."""
    import hashlib

    h = hashlib.sha256()
    for f in KERNEL_SOURCES:
        with open(os.path.join(ROOT, "q-vti_amd", "csrc", f), "rb") as fh:
            h.update(fh.read())
    return h.hexdigest()[:16]


def measured_traffic(name, kernel):
    """HBM bytes per launch of `kernel` on workload `name` from the committed
    rocprofv3 FETCH_SIZE / WRITE_SIZE summary (scaled by the copy probe), or
    None when no summary covers it or the kernel sources changed since."""
    path = os.path.join(ROOT, "profiles", "%s_traffic.json" % name)
    try:
        d = json.load(open(path))
        if d.get("kernel_sources_sha16") != kernel_source_hash():
            return None
        k = d["kernels"][kernel]
        return round(k["fetch_bytes"] + k["write_bytes"])
    except (OSError, KeyError, ValueError):
        return None


# Kernels of one time step in launch order (the engine's enqueue_step): the
# name each dispatch carries in a rocprofv3 trace, and the class it times.
STEP_SIGNATURE = (("k2_deriv_strided", "vel_x"), ("k2_deriv_strided", "vel_y"), ("k2_vel_z", "vel_z_fwd"),
                  ("k2_cfft_strided", "y_fwd"), ("k2_xpass", "xpass"), ("k2_cfft_strided", "y_inv"),
                  ("k2_zrows", "z_inv"), ("k2_deriv_strided", "str_x"), ("k2_deriv_strided", "str_y"),
                  ("k2_stress_z", "str_z"), ("k_record", "record"))


def _rocprof_rows(d, suffix):
    """Rows of every rocprofv3 CSV under directory d whose name ends in suffix."""
    import csv
    import glob

    rows = []
    for f in sorted(glob.glob(os.path.join(d, "**", "*" + suffix), recursive=True)):
        with open(f, newline="") as fh:
            rows += list(csv.DictReader(fh))
    if not rows:
        raise SystemExit("no *%s under %s" % (suffix, d))
    return rows


def _dispatches(rows, counter=None):
    """[(dispatch id, kernel name, value)] in dispatch order: the duration in ns
    for kernel-trace rows, the summed counter value for counter rows."""
    per = {}
    for r in rows:
        did = int(r["Dispatch_Id"])
        if counter is None:
            per[did] = (r["Kernel_Name"], int(r["End_Timestamp"]) - int(r["Start_Timestamp"]))
        elif r["Counter_Name"] == counter:
            name, v = per.get(did, (r["Kernel_Name"], 0.0))
            per[did] = (name, v + float(r["Counter_Value"]))
    return [(d, n, v) for d, (n, v) in sorted(per.items())]


def _step_classes(disp):
    """Per class, the values of the dispatches that form complete time steps
    (runs that match STEP_SIGNATURE; tuning launches outside steps are left
    out)."""
    out = {c: [] for _, c in STEP_SIGNATURE}
    sig = [k for k, _ in STEP_SIGNATURE]
    i = 0
    while i + len(sig) <= len(disp):
        names = [disp[i + j][1] for j in range(len(sig))]
        if all(k in n for k, n in zip(sig, names)):
            for j, (_, c) in enumerate(STEP_SIGNATURE):
                out[c].append(disp[i + j][2])
            i += len(sig)
        else:
            i += 1
    return out


def summarize_counters(workload, trace_dir, fetch_dir, write_dir, copy_fetch_dir, copy_write_dir, bench_json):
    """Turns the rocprofv3 outputs of one workload into the committed counter
    summary profiles/<workload>_traffic.json (read back by measured_traffic)
    and prints a markdown table.  trace_dir: --kernel-trace --stats run of
    bench.py; fetch_dir / write_dir: --pmc FETCH_SIZE / WRITE_SIZE runs of
    bench.py; copy_*_dir: the same counters over `bench.py --copy-probe`;
    bench_json: the JSON line of the traced run (algorithmic bytes)."""
    import statistics

    probe = 1024 << 20
    copies = 4  # copy_probe: one warm-up copy and 3 timed ones

    def copy_kib(d, counter):
        v = [x for _, n, x in _dispatches(_rocprof_rows(d, "counter_collection.csv"), counter) if "opyBuffer" in n]
        if len(v) < copies:
            raise SystemExit("copy probe: %d copy dispatches under %s" % (len(v), d))
        return sum(v[-copies:]) / copies

    fetch_factor = probe / (1024.0 * copy_kib(copy_fetch_dir, "FETCH_SIZE"))
    write_factor = probe / (1024.0 * copy_kib(copy_write_dir, "WRITE_SIZE"))
    times = _step_classes(_dispatches(_rocprof_rows(trace_dir, "kernel_trace.csv")))
    fetch = _step_classes(_dispatches(_rocprof_rows(fetch_dir, "counter_collection.csv"), "FETCH_SIZE"))
    write = _step_classes(_dispatches(_rocprof_rows(write_dir, "counter_collection.csv"), "WRITE_SIZE"))
    line = json.loads(open(bench_json).read().strip().splitlines()[-1])
    kernels = {}
    for _, c in STEP_SIGNATURE:
        if not times[c] or not fetch[c] or not write[c] or c not in line["kernels"]:
            continue
        k = line["kernels"][c]
        kernels[c] = {"fetch_bytes": 1024.0 * statistics.mean(fetch[c]) * fetch_factor,
                      "write_bytes": 1024.0 * statistics.mean(write[c]) * write_factor,
                      "algorithmic_bytes": k["GBps"] * k["avg_us"] * 1e3,
                      "rocprof_avg_us": statistics.mean(times[c]) / 1e3,
                      "bench_event_avg_us": k["avg_us"],
                      "steps_traced": len(times[c]), "steps_counted": min(len(fetch[c]), len(write[c]))}
    out = {"workload": workload, "kernel_sources_sha16": kernel_source_hash(), "fetch_factor": fetch_factor,
           "write_factor": write_factor, "kernels": kernels}
    with open(os.path.join(ROOT, "profiles", "%s_traffic.json" % workload), "w") as fh:
        json.dump(out, fh, indent=1)
    print("FETCH_SIZE x %.4f, WRITE_SIZE x %.4f (copy probe, %d MiB each way)\n" % (fetch_factor, write_factor,
                                                                                  probe >> 20))
    print("| kernel class | rocprof avg us | bench HIP-event avg us | algorithmic MB | fetch MB | write MB "
          "| traffic / algorithmic | algorithmic GB/s | of 8 TB/s |")
    print("|---|---|---|---|---|---|---|---|---|")
    tot = 0.0
    for c, k in kernels.items():
        tot += k["rocprof_avg_us"]
        gbs = k["algorithmic_bytes"] / (k["rocprof_avg_us"] * 1e3)
        print("| %s | %.1f | %.1f | %.1f | %.1f | %.1f | %.3f | %.0f | %.3f |" % (
            c, k["rocprof_avg_us"], k["bench_event_avg_us"], k["algorithmic_bytes"] / 1e6, k["fetch_bytes"] / 1e6,
            k["write_bytes"] / 1e6, (k["fetch_bytes"] + k["write_bytes"]) / max(k["algorithmic_bytes"], 1.0), gbs,
            gbs / HBM_PEAK_GBS))
    print("\nsum of per-step kernel time (rocprof): %.0f us; bench wall time per step %.0f us (%.4f Gcell-steps/s)"
          % (tot, line["ms_per_step"] * 1e3, line["value"]))


def summarize_pmc(dirs):
    """Per step kernel class, the mean per-dispatch value of every counter in
    the rocprofv3 --pmc runs under dirs (one pass each), as a markdown table."""
    import statistics

    table = {}
    for d in dirs:
        rows = _rocprof_rows(d, "counter_collection.csv")
        for counter in sorted({r["Counter_Name"] for r in rows}):
            per = _step_classes(_dispatches(rows, counter))
            table[counter] = {c: statistics.mean(v) for c, v in per.items() if v}
    classes = [c for _, c in STEP_SIGNATURE if any(c in t for t in table.values())]
    print("| counter | " + " | ".join(classes) + " |")
    print("|---|" + "---|" * len(classes))
    for counter, t in table.items():
        print("| %s | " % counter + " | ".join("%.4g" % t[c] if c in t else "" for c in classes) + " |")


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


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
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": "%s FP32 oracle, %d-step run(s), per-step time from the difference of a 3-step and a "
                      "1-step run (setup excluded)" % (name, steps)}


def surface_shots(g, n, seed):
    """n sources at seeded uniform (x, y) inside the C-PML frame, z = pml + 1,
    each with the workload's receiver line."""
    rng = np.random.default_rng(seed)
    p = g["pml"]
    xs = rng.integers(p, g["nx"] - p, n)
    ys = rng.integers(p, g["ny"] - p, n)
    rx = np.arange(g["nx"])
    return [qvti.make_shot(qvti.SRC_FZ, (int(x), int(y), p + 1), rx, np.full(g["nx"], g["ny"] // 2),
                           np.full(g["nx"], p + 1)) for x, y in zip(xs, ys)]


def time_steps(e, nt, rik, shots, steps, warmup):
    e.begin(nt, rik, shots)
    e.step(warmup)
    e.sync()
    t0 = time.perf_counter()
    e.step(steps)
    e.sync()
    return time.perf_counter() - t0


def run_batched(steps, warmup, nshots=32, nt_full=2000, workload="vti160"):
    """32 shots of one model (default: the 160x160x128 VTI workload): the
    plain path (one shot per launch sequence) against batched launches (W
    shots per kernel launch, W picked by the library's tuner from silent
    timing runs of widths 1, 2, 4, ... 32).  shots/s is for nt_full-step
    shots, from the timed per-step wall time of each path.
    At the chosen width a second region launches the step's kernels directly
    with HIP events around each launch: the sum of kernel times per step
    against the HIP-graph wall time per step is the launch overhead a
    persistent kernel could remove."""
    g, m, _ = make_workload(workload)
    shots = surface_shots(g, nshots, 11)
    nt = warmup + 2 * steps
    rik = qvti.ricker(nt, g["dt"], g["f0"])
    with engine(g) as e:
        e.set_model_dict(m)
        e.set_probe(0)
        plain_wall = time_steps(e, nt, rik, shots[:1], steps, warmup)
    plain_step = plain_wall / steps
    with engine(g, max_shots=nshots) as e:
        e.set_model_dict(m)
        e.set_probe(0)
        # the library's batch-width tuner (qvti_batch_width) picks the shots per launch
        wbest, lib_trials = e.batch_width(shots)
        wall = time_steps(e, nt, rik, shots[:wbest], steps, warmup)
        per_shot_step = wall / steps / wbest
        e.enable_kernel_timing(steps)
        e.step(steps)
        e.sync()
        kt = e.kernel_times()
        kernel_sum = (wbest, sum(v["total_ms"] for v in kt.values()) / steps / 1e3, wall / steps)
        per_kernel = {n: {"avg_us": round(1e3 * v["total_ms"] / v["launches"], 2),
                          "GBps": round(v["bytes_per_launch"] / (v["total_ms"] / v["launches"] * 1e-3) / 1e9, 1)}
                      for n, v in kt.items() if v["launches"] > 0}
    trials = {w: ms * 1e-3 for w, ms in lib_trials.items()}
    N = g["nx"] * g["ny"] * g["nz"]
    balg = b_alg(g["nx"], g["ny"], g["nz"], g["pml"], g["vti"])
    plain_sps = 1.0 / (plain_step * nt_full)
    batched_sps = 1.0 / (per_shot_step * nt_full)
    gcs_plain, gcs_batch = N / plain_step / 1e9, N / per_shot_step / 1e9
    wl = "%dx%dx%d %s" % (g["nx"], g["ny"], g["nz"], "VTI" if g["vti"] else "isotropic")
    out = {
        "metric": "batched shots/s (%s, %d-step shots)" % (wl, nt_full),
        "value": round(batched_sps, 4), "unit": "shots/s", "steps": steps, "warmup": warmup,
        "ms_per_step": round(per_shot_step * wbest * 1e3, 4),
        "config": {"workload": "%d shots on one %s model, pml %d, dx %.1f m, dt %.0e s, sources seeded on the "
                               "surface grid" % (nshots, wl, g["pml"], g["dx"], g["dt"]),
                   "cells": N, "parallelism": "1 GPU, %d shots per launch" % wbest},
        "shots_per_launch": wbest,
        "plain_shots_per_s": round(plain_sps, 4),
        "speedup_vs_plain": round(batched_sps / plain_sps, 3),
        "gcell_steps_per_s": {"plain": round(gcs_plain, 3), "batched": round(gcs_batch, 3)},
        "roofline_frac": {"b_alg_bytes_per_cell_step": round(balg, 1), "plain": round(gcs_plain * balg / HBM_PEAK_GBS, 4),
                          "batched": round(gcs_batch * balg / HBM_PEAK_GBS, 4)},
        "width_trials_ms_per_shot_step": {str(w): round(v * 1e3, 4) for w, v in sorted(trials.items())},
        "width_chosen_by": "qvti_batch_width (library tuner, silent timing runs)",
    }
    w, ks, wall = kernel_sum
    out["graph_step"] = {"shots": w, "kernel_sum_ms": round(ks * 1e3, 4), "graph_wall_ms": round(wall * 1e3, 4),
                         "launch_overhead_frac": round(1.0 - ks / wall, 4)}
    out["kernels"] = per_kernel
    return out


def stream_copy(nbytes=2 << 30, reps=10):
    """Measured device-to-device stream copy: hipMemcpyAsync (the HIP
    runtime's blit kernel) of 2 GiB, HIP-event timed on one stream; bytes
    read + written per second."""
    import hiprt

    s, bps = hiprt.stream_copy(nbytes, reps)
    gbs = bps / 1e9
    return {"GBps": round(gbs, 1), "frac_of_peak": round(gbs / HBM_PEAK_GBS, 4), "bytes_each_way": nbytes,
            "reps": reps, "us_per_copy": round(s * 1e6, 1), "kernel": "HIP runtime device-to-device copy"}


SERVE_GRIDS = [(96, 96, 96), (128, 128, 128), (160, 160, 128), (200, 200, 100)]


def serve_requests(n, seed=5, nt_range=(1000, 3000), rate=4.0):
    """Seeded Poisson stream: grid from SERVE_GRIDS, isotropic or VTI with
    probability 0.5, nt uniform in nt_range, Q in [20, 200], vp in
    [2000, 4000], dx 12.5 m, dt 0.8 ms, 20-cell C-PML."""
    rng = np.random.default_rng(seed)
    t, reqs = 0.0, []
    for i in range(n):
        t += rng.exponential(1.0 / rate)
        nx, ny, nz = SERVE_GRIDS[rng.integers(len(SERVE_GRIDS))]
        vti = bool(rng.random() < 0.5)
        nt = int(rng.integers(nt_range[0], nt_range[1] + 1))
        reqs.append(dict(t=t, shape=(nx, ny, nz), vti=vti, nt=nt, seed=int(rng.integers(1 << 30))))
    return reqs


def serve_inputs(r):
    nx, ny, nz = r["shape"]
    pml, d, dt, f0 = 20, 12.5, 8e-4, 20.0
    m = qvti.make_model(2, r["seed"], (nx, ny, nz), [2000, 4000, 0.2, 20, 200, 8, 0.05], vti=r["vti"])
    m["gp"] = qvti.gamma_from_q(m["qp"])
    m["gs"] = qvti.gamma_from_q(m["qs"])
    if not r["vti"]:
        m.pop("eps", None)
        m.pop("delta", None)
    grid = qvti.Grid(nx, ny, nz, pml, d, d, d, dt, f0, float(np.float32(2 * PI * f0)), int(r["vti"]))
    rx = np.arange(nx)
    shot = qvti.make_shot(qvti.SRC_FZ, (nx // 2, ny // 2, pml + 1), rx, np.full(nx, ny // 2), np.full(nx, pml + 1))
    rik = qvti.ricker(r["nt"], dt, f0)
    return grid, m, shot, rik


def run_served(nreq, max_active=4, slice_steps=50):
    """Continuous batching against plain sequential execution of the same
    seeded Poisson stream.  Sequential latencies come from each request's
    measured plain service time in a FIFO single-server queue with the same
    arrival times."""
    reqs = serve_requests(nreq)
    inputs = [serve_inputs(r) for r in reqs]
    # plain sequential service times (context per request, as a caller of qvti_forward would)
    service = []
    for r, (grid, m, shot, rik) in zip(reqs, inputs):
        t0 = time.perf_counter()
        with qvti.Engine(grid.nx, grid.ny, grid.nz, grid.pml, grid.dx, grid.dy, grid.dz, grid.dt, grid.f0, grid.w0,
                         vti=bool(grid.vti)) as e:
            e.set_model_dict(m)
            e.forward(r["nt"], rik, shot)
        service.append(time.perf_counter() - t0)
    free, seq_lat = 0.0, []
    for r, s in zip(reqs, service):
        start = max(free, r["t"])
        free = start + s
        seq_lat.append(free - r["t"])
    seq_span = free - reqs[0]["t"]
    # served
    budget = 200 * (1 << 30)
    lat = []
    with qvti.Server(budget, max_active=max_active, steps_per_slice=slice_steps) as srv:
        t0 = srv.clock()
        ids = []
        for r, (grid, m, shot, rik) in zip(reqs, inputs):
            delay = t0 + r["t"] - srv.clock()
            if delay > 0:
                time.sleep(delay)
            ids.append(srv.submit(grid, r["nt"], rik, m, shot))
        for i in ids:
            rc = srv.wait(i, 900.0)
            if rc != 0:
                raise RuntimeError("request %d ended with %d" % (i, rc))
        infos = [srv.info(i)[1] for i in ids]
        lat = [inf.t_done_s - inf.t_submit_s for inf in infos]
        span = max(inf.t_done_s for inf in infos) - min(inf.t_submit_s for inf in infos)
    cells = [int(np.prod(r["shape"])) * r["nt"] for r in reqs]
    return {
        "metric": "served shots/s and p50/p99 latency (seeded Poisson stream, 4 req/s offered)",
        "value": round(nreq / span, 4), "unit": "shots/s", "n_gpus": 1, "steps": int(sum(r["nt"] for r in reqs)),
        "warmup": 0, "ms_per_step": None, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "fp32", "data": "synthetic (seeded smooth random models)",
        "config": {"workload": "%d requests, grids 96^3/128^3/160x160x128/200x200x100, iso or VTI, nt in "
                               "[1000,3000], Q in [20,200], dx 12.5 m, dt 0.8 ms" % nreq,
                   "parallelism": "1 GPU, up to %d requests in flight, %d-step slices" % (max_active, slice_steps)},
        "latency_s": {"p50": round(float(np.percentile(lat, 50)), 3), "p99": round(float(np.percentile(lat, 99)), 3)},
        "sequential": {"shots_per_s": round(nreq / seq_span, 4),
                       "p50": round(float(np.percentile(seq_lat, 50)), 3),
                       "p99": round(float(np.percentile(seq_lat, 99)), 3)},
        "speedup_vs_sequential": round(seq_span / span, 3),
        "offered_gcell_steps_per_s": round(sum(cells) / (reqs[-1]["t"] - reqs[0]["t"]) / 1e9, 3),
        "served_gcell_steps_per_s": round(sum(cells) / span / 1e9, 3),
    }


def copy_probe(mib=1024, reps=3):
    """A plain device copy of known size (the HIP runtime's device-to-device
    blit kernel): the reference point that scales hardware FETCH_SIZE /
    WRITE_SIZE counters to bytes."""
    import hiprt

    hiprt.stream_copy(mib << 20, reps)
    return {"copy_probe_bytes_each_way": mib << 20, "reps": reps}


LEG_DROP = ("n_gpus", "higher_is_better", "scaling", "vs_baseline", "dtype", "data")


def leg(fn, *a, **k):
    """One extra leg of the default line: its result without the keys the
    headline already states, or the error it raised (the headline stands on
    its own)."""
    t0 = time.perf_counter()
    try:
        out = {key: v for key, v in fn(*a, **k).items() if key not in LEG_DROP}
    except Exception as ex:
        out = {"error": "%s: %s" % (type(ex).__name__, ex)}
    out["leg_wall_s"] = round(time.perf_counter() - t0, 1)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", default="vti256", choices=sorted(WORKLOADS))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--headline-only", action="store_true",
                    help="skip the batched, served and stream-copy legs of the default line")
    ap.add_argument("--batched", action="store_true", help="32-shot batched leg instead (--workload, default vti160)")
    ap.add_argument("--served", type=int, default=0, metavar="NREQ", help="serving leg with NREQ requests instead")
    ap.add_argument("--copy-probe", action="store_true", help="only run a 1 GiB device copy (counter calibration)")
    ap.add_argument("--tiles", type=int, default=1, metavar="MODE",
                    help="tile mode of the strided kernels (qvti_set_autotune; default 1 = autotuned)")
    ap.add_argument("--kspace", type=float, default=0.0, metavar="VREF",
                    help="time the opt-in k-space corrected scheme (reference speed VREF m/s) instead")
    ap.add_argument("--summarize-counters", nargs=6, metavar=("TRACE", "FETCH", "WRITE", "COPY_FETCH", "COPY_WRITE",
                                                              "BENCH_JSON"),
                    help="write profiles/<workload>_traffic.json from rocprofv3 outputs (no GPU needed)")
    ap.add_argument("--summarize-pmc", nargs="+", metavar="DIR",
                    help="per-class table of the counters of rocprofv3 --pmc runs of bench.py (no GPU needed)")
    args = ap.parse_args()
    if args.summarize_pmc:
        summarize_pmc(args.summarize_pmc)
        return
    if args.summarize_counters:
        summarize_counters(args.workload, *args.summarize_counters)
        return
    if args.copy_probe:
        print(json.dumps(copy_probe()), flush=True)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU only (shots shard across GPUs with no communication); "
              "--gpus %d is not supported" % args.gpus, file=sys.stderr)
        sys.exit(2)
    if args.batched:
        wl = args.workload if args.workload != "vti256" else "vti160"
        nt_full = 500 if wl == "iso64" else (3000 if wl in ("vti256", "vti400") else 2000)
        print(json.dumps(run_batched(args.steps, args.warmup, workload=wl, nt_full=nt_full)), flush=True)
        return
    if args.served:
        print(json.dumps(run_served(args.served)), flush=True)
        return
    name = args.workload
    g, wall, wall_timed, kt, setup_s, begin_s, tiles, blow_up = run_plain(name, args.steps, args.warmup, args.kspace, args.tiles)
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
        # B_alg is the plain scheme's; the k-space mode moves more bytes per step
        "roofline": {"bound": "hbm", "kernel": best, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                     "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                     "algorithmic_bytes": round(k["bytes_per_launch"]), "avg_launch_us": round(avg_ms * 1e3, 2),
                     "traffic": None if args.kspace > 0 else measured_traffic(name, best)},
        "step_roofline": {"b_alg_bytes_per_cell_step": round(balg, 1),
                          "achieved_GBps": round(balg * gcs, 1),
                          "frac": round(balg * gcs / HBM_PEAK_GBS, 4)},
        "kernels": per_kernel,
        "ms_per_step_event_timed": round(wall_timed / args.steps * 1e3, 4),
        "setup_s": round(setup_s, 3),
        "begin_s": round(begin_s, 3),
        "tile_pencils": tiles,
    }
    if args.kspace > 0:
        out["config"]["kspace_vref"] = args.kspace
    if args.tiles != 1:
        out["config"]["tile_mode"] = args.tiles
    if blow_up is not None:
        out["stability"] = ("the blow-up probe fired by step %d: dt %.0e s is beyond the scheme's stability limit "
                            "for this model (the work per step is unchanged)" % (blow_up, g["dt"]))
    if not args.headline_only and args.kspace == 0 and name == "vti256":
        out["batched"] = leg(run_batched, 30, 5)
        out["served"] = leg(run_served, 16)
        out["stream_copy"] = leg(stream_copy)
    if not args.no_cpu_baseline:
        try:
            out["cpu_baseline"] = run_cpu_baseline(name)
        except Exception as ex:  # the GPU number stands on its own
            out["cpu_baseline"] = {"error": str(ex)}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
