"""Build the in-tree shared libraries.

* ``_build/libpureibm_gpu.so``  HIP kernels + C-ABI (hipcc, gfx950)
* ``_build/libpib_host.so``     Fortran host: namelists, set-up (amdflang)
* ``_build/pureibm_gpu``        Fortran driver program linked to both

Run ``python -m pureibm_ps2d_fd1d_amd.build`` or ``__graft_entry__.build()``.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
FSRC = os.path.join(HERE, "fsrc")
BUILD = os.path.join(HERE, "_build")

HIP_FLAGS = [
    "--offload-arch=gfx950", "-O3", "-std=c++17", "-fPIC", "-shared",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result",
]


def _run(cmd, cwd=None):
    print("+", " ".join(cmd), flush=True)
    subprocess.check_call(cmd, cwd=cwd)


def hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flang():
    for c in ("/opt/rocm/bin/amdflang", shutil.which("amdflang"), shutil.which("flang")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("amdflang not found")


def build_gpu():
    """Each translation unit compiled on its own process (they build in
    parallel), then linked into the shared library."""
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, "libpureibm_gpu.so")
    srcs = ["pib_gpu.hip", "pib_dem.hip", "pib_adjoint.hip"]
    objs = [os.path.join(BUILD, os.path.splitext(f)[0] + ".o") for f in srcs]
    flags = [f for f in HIP_FLAGS if f != "-shared"] + ["-I", os.path.join(ROOT, "include")]
    procs = []
    for src, obj in zip(srcs, objs):
        cmd = [hipcc()] + flags + ["-c", os.path.join(CSRC, src), "-o", obj]
        print("+", " ".join(cmd), flush=True)
        procs.append(subprocess.Popen(cmd))
    if any(p.wait() != 0 for p in procs):
        raise subprocess.CalledProcessError(1, "hipcc")
    _run([hipcc(), "--offload-arch=gfx950", "-shared", "-fPIC"] + objs + ["-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_host():
    os.makedirs(BUILD, exist_ok=True)
    out = os.path.join(BUILD, "libpib_host.so")
    srcs = [os.path.join(FSRC, f) for f in ("pib_types.f90", "pib_host.f90")]
    _run([flang(), "-O2", "-fPIC", "-shared", "-Wno-real-constant-widening", "-module-dir", BUILD]
         + srcs + ["-o", out + ".tmp"])
    os.replace(out + ".tmp", out)
    return out


def build_driver():
    out = os.path.join(BUILD, "pureibm_gpu")
    srcs = [os.path.join(FSRC, f) for f in ("pib_types.f90", "pib_host.f90", "pib_main.f90")]
    if not os.path.exists(srcs[-1]):
        return None
    _run([flang(), "-O2", "-Wno-real-constant-widening", "-module-dir", BUILD] + srcs
         + ["-o", out + ".tmp", "-L" + BUILD, "-lpureibm_gpu", "-Wl,-rpath,$ORIGIN"])
    os.replace(out + ".tmp", out)
    return out


def build_all():
    build_host()
    build_gpu()
    build_driver()


if __name__ == "__main__":
    what = sys.argv[1:] or ["all"]
    for w in what:
        {"all": build_all, "gpu": build_gpu, "host": build_host, "driver": build_driver}[w]()
