"""Writes the full-size parity fixtures: receiver gathers of the FP32 and FP64
CPU oracle for the cases of tests/fullsize_cases.py.

  fullsize_vti256.npz   256^3 smooth random VTI, 200 steps of 0.5 ms
  fullsize_vti400.npz   400x400x200 smooth random VTI, 160 steps of 0.5 ms
  batch_vti160.npz      shot 0 of the 32-shot 160x160x128 batch, 200 steps of 0.5 ms
  serve_grids.npz       one request on each serving grid (96^3 iso, 128^3 VTI,
                        160x160x128 iso, 200x200x100 VTI), 150 steps of 0.8 ms

Each file holds rec_f32 / rec_f64 [3][nt][R] (vx, vy, vz), the wavelet, the
receiver cells and a checksum of the regenerated model.  Run once in the build
container (about two hours on 6 cores; the 400x400x200 FP64 run holds ~37 GB):
    python tests/golden/make_fullsize_golden.py [case ...]
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path[:0] = [os.path.join(ROOT, "oracle"), os.path.join(ROOT, "q-vti_amd"), os.path.join(ROOT, "tests")]
import oracle as O  # noqa: E402
import fullsize_cases as FC  # noqa: E402

THREADS = int(os.environ.get("QVTI_ORACLE_THREADS", "6"))


def run_both(g, m, shot, nt, rik):
    s, model = FC.oracle_inputs(g, m, shot, nt, rik)
    t = time.time()
    r32 = O.run(s, model, "f32", nthreads=THREADS)["records"]
    print("  f32 %.0fs" % (time.time() - t), flush=True)
    t = time.time()
    r64 = O.run(s, model, "f64", nthreads=THREADS)["records"]
    print("  f64 %.0fs" % (time.time() - t), flush=True)
    return r32.astype(np.float32), r64.astype(np.float64)


def make_fullsize(name):
    g, m, shot = FC.fullsize_case(name)
    print(name, flush=True)
    r32, r64 = run_both(g, m, shot, shot["nt"], shot["rik"])
    np.savez_compressed(os.path.join(HERE, "fullsize_%s.npz" % name), rec_f32=r32, rec_f64=r64, rik=shot["rik"],
                        rx=shot["rx"], ry=shot["ry"], rz=shot["rz"], src=np.array(shot["src"]),
                        model_sum=FC.model_checksum(m))


def make_batch():
    g, m, shots = FC.batch_case()
    print("batch_vti160 shot 0", flush=True)
    rik = FC.wavelet(FC.BATCH_NT)
    r32, r64 = run_both(g, m, shots[0], FC.BATCH_NT, rik)
    np.savez_compressed(os.path.join(HERE, "batch_vti160.npz"), rec_f32=r32, rec_f64=r64, rik=rik,
                        src=np.array(shots[0]["src"]), rx=shots[0]["rx"], model_sum=FC.model_checksum(m))


def make_serve():
    out = {}
    for k in range(len(FC.SERVE_CASES)):
        g, m, shot = FC.serve_case(k)
        print("serve %d %s" % (k, FC.SERVE_CASES[k]), flush=True)
        r32, r64 = run_both(g, m, shot, shot["nt"], shot["rik"])
        out["rec_f32_%d" % k] = r32
        out["rec_f64_%d" % k] = r64
        out["model_sum_%d" % k] = FC.model_checksum(m)
        out["rik_%d" % k] = shot["rik"]
    np.savez_compressed(os.path.join(HERE, "serve_grids.npz"), **out)


if __name__ == "__main__":
    todo = sys.argv[1:] or ["serve", "batch", "vti256", "vti400"]
    for name in todo:
        if name == "serve":
            make_serve()
        elif name == "batch":
            make_batch()
        else:
            make_fullsize(name)
