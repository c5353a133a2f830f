"""Build recipe of the engine library (q-vti_amd/libqvti.so) for gfx950.

Run `python q-vti_amd/build.py` (or __graft_entry__.build()).  Always
recompiles.  Every HIP translation unit gets the packed-FP32 target feature
switched off; the host pass of hipcc warns that it does not know the
feature, which is harmless.
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libqvti.so")

HIP_SOURCES = ["qvti_engine.hip", "qvti_serve.hip", "qvti_rtm.hip"]
# qvti_kernels.hip is compiled once per build group of axis lengths
# (QVTI_NGROUPS in qvti_lengths.h), the groups in parallel
KERNEL_SOURCE = "qvti_kernels.hip"
NGROUPS = 8
CPP_SOURCES = ["qvti_host.cpp", "qvti_error.cpp"]
ARCH = os.environ.get("QVTI_ARCH", "gfx950")


def hipcc():
    p = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return p if os.path.exists(p) else "hipcc"


def build(verbose=False, out=None, defines=()):
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC] + ["-D" + d for d in defines]
    common = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result"]
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    hip_flags = ["-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"]
    jobs = []
    for src in HIP_SOURCES:
        obj = os.path.join(bdir, src + ".o")
        jobs.append((obj, [hipcc(), "-x", "hip", "--offload-arch=" + ARCH, "-c", os.path.join(CSRC, src), "-o", obj]
                     + hip_flags + common + inc))
    for g in range(NGROUPS):
        obj = os.path.join(bdir, "%s.g%d.o" % (KERNEL_SOURCE, g))
        jobs.append((obj, [hipcc(), "-x", "hip", "--offload-arch=" + ARCH, "-c", os.path.join(CSRC, KERNEL_SOURCE),
                           "-o", obj, "-DQVTI_GROUP=%d" % g] + hip_flags + common + inc))
    for src in CPP_SOURCES:
        obj = os.path.join(bdir, src + ".o")
        jobs.append((obj, [hipcc(), "-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj] + common + inc))
    _run_parallel([cmd for _, cmd in jobs], verbose)
    objs = [obj for obj, _ in jobs]
    lib = out or LIB
    tmp = lib + ".tmp"
    _run([hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"], verbose)
    os.replace(tmp, lib)
    if out is None:  # the command-line host links the in-tree library
        build_host(lib, verbose)
    return lib


HOST = os.path.join(HERE, "qvti_shots")


def build_host(lib=LIB, verbose=False):
    """The C++ host driver (host/qvti_shots.cpp) linked against the library;
    it finds libqvti.so next to itself at run time."""
    tmp = HOST + ".tmp"
    _run([hipcc(), "-x", "c++", "-O2", "-std=c++17", "-Wall", "-I", os.path.join(ROOT, "include"),
          os.path.join(HERE, "host", "qvti_shots.cpp"), "-o", tmp, "-L", os.path.dirname(lib), "-lqvti",
          "-Wl,-rpath,$ORIGIN"], verbose)
    os.replace(tmp, HOST)
    return HOST


def _run_parallel(cmds, verbose, jobs=None):
    """Run the compile commands concurrently (at most `jobs`, default: the
    CPU count capped at 16); raise on the first failure after all finished."""
    jobs = jobs or min(16, os.cpu_count() or 1)
    procs, pending, failed = [], list(cmds), []
    while pending or procs:
        while pending and len(procs) < jobs:
            cmd = pending.pop(0)
            if verbose:
                print(" ".join(cmd), flush=True)
            procs.append((cmd, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)))
        cmd, p = procs.pop(0)
        out, _ = p.communicate()
        if p.returncode != 0:
            failed.append((cmd, out))
        elif verbose and out:
            print("\n".join(l for l in out.splitlines() if "packed-fp32-ops" not in l and l.strip()))
    if failed:
        for cmd, out in failed:
            sys.stderr.write(out)
        raise RuntimeError("build step failed: " + " ".join(failed[0][0]))


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    out = "\n".join(l for l in r.stdout.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        sys.stderr.write(r.stdout)
        raise RuntimeError("build step failed: " + " ".join(cmd))
    if out and verbose:
        print(out)


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
