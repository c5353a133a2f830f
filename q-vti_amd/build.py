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

HIP_SOURCES = ["qvti_engine.hip", "qvti_serve.hip"]
CPP_SOURCES = ["qvti_host.cpp", "qvti_error.cpp"]
ARCH = os.environ.get("QVTI_ARCH", "gfx950")


def hipcc():
    p = os.path.join(os.environ.get("ROCM_PATH", "/opt/rocm"), "bin", "hipcc")
    return p if os.path.exists(p) else "hipcc"


def build(verbose=False, out=None, defines=()):
    inc = ["-I", os.path.join(ROOT, "include"), "-I", CSRC] + ["-D" + d for d in defines]
    common = ["-O3", "-fPIC", "-std=c++17", "-Wall", "-Wno-unused-function", "-Wno-unused-value", "-Wno-unused-result"]
    objs = []
    bdir = os.path.join(HERE, "_build")
    os.makedirs(bdir, exist_ok=True)
    for src in HIP_SOURCES:
        obj = os.path.join(bdir, src + ".o")
        cmd = [hipcc(), "-x", "hip", "--offload-arch=" + ARCH, "-c", os.path.join(CSRC, src), "-o", obj,
               "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-munsafe-fp-atomics"] + common + inc
        _run(cmd, verbose)
        objs.append(obj)
    for src in CPP_SOURCES:
        obj = os.path.join(bdir, src + ".o")
        cmd = [hipcc(), "-x", "c++", "-c", os.path.join(CSRC, src), "-o", obj] + common + inc
        _run(cmd, verbose)
        objs.append(obj)
    lib = out or LIB
    tmp = lib + ".tmp"
    _run([hipcc(), "-shared", "-fPIC", "--offload-arch=" + ARCH, "-o", tmp] + objs + ["-lpthread"], verbose)
    os.replace(tmp, lib)
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
