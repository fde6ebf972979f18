"""TEST INFRASTRUCTURE ONLY: ctypes loader for the CPU oracle (pib_oracle.c).

Used by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg as
the checker.  The product never imports this module.
"""
import ctypes
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "_build", "liboracle.so")
SRC = os.path.join(HERE, "pib_oracle.c")
SRCS = [SRC, os.path.join(HERE, "pib_dem_oracle.c")]

_lib = None


def build(force=False):
    """Compile the oracle with gcc (FP contraction off, so each op rounds as
    written; OpenMP for the independent plane and mode loops)."""
    os.makedirs(os.path.dirname(LIB), exist_ok=True)
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < max(map(os.path.getmtime, SRCS)):
        subprocess.check_call(["gcc", "-O2", "-ffp-contract=off", "-fopenmp", "-fPIC", "-shared",
                               "-o", LIB + ".tmp"] + SRCS + ["-lm"])
        os.replace(LIB + ".tmp", LIB)
    return LIB


def load():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        import sys
        sys.path.insert(0, os.path.dirname(HERE))
        from pureibm_ps2d_fd1d_amd import _abi
        lib = ctypes.CDLL(LIB)
        _abi.bind(lib, "orc_", _abi._STEP_SIGS)
        _abi.bind(lib, "orc_", _abi._DEM_SIGS)
        lib.orc_dem_neighbours.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_int)]
        lib.orc_dem_neighbours.restype = ctypes.c_int
        PD = ctypes.POINTER(ctypes.c_double)
        lib.orc_ff2rc.argtypes = [ctypes.c_int, ctypes.c_int, PD, PD]
        lib.orc_ff2rc.restype = None
        lib.orc_ff2cr.argtypes = [ctypes.c_int, ctypes.c_int, PD, PD]
        lib.orc_ff2cr.restype = None
        lib.orc_fft1d.argtypes = [ctypes.c_int, ctypes.c_int, PD, PD]
        lib.orc_fft1d.restype = None
        lib.orc_tridag.argtypes = [PD, PD, PD, PD, PD, ctypes.c_int]
        lib.orc_tridag.restype = ctypes.c_int
        lib.orc_cyclic_solve.argtypes = [PD, PD, PD, PD, PD, ctypes.c_int]
        lib.orc_cyclic_solve.restype = ctypes.c_int
        lib.orc_quadgener.argtypes = [ctypes.c_double, ctypes.c_double, ctypes.c_double,
                                      ctypes.c_double, ctypes.c_int, PD, PD]
        lib.orc_quadgener.restype = ctypes.c_int
        lib.orc_set_threads.argtypes = [ctypes.c_int]
        lib.orc_set_threads.restype = None
        lib.orc_get_threads.argtypes = []
        lib.orc_get_threads.restype = ctypes.c_int
        lib.orc_pressure_modes.argtypes = [ctypes.c_void_p, ctypes.c_int]
        lib.orc_pressure_modes.restype = ctypes.c_int
        lib.orc_get_if_rev.argtypes = [ctypes.c_void_p, ctypes.POINTER(ctypes.c_ubyte)]
        lib.orc_get_if_rev.restype = ctypes.c_int
        _lib = lib
    return _lib


def _p(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def ff2rc(a):
    """a: real (my, mz) -> complex (my/2+1, mz), forward scaled by 1/(my*mz)."""
    a = np.asfortranarray(a, dtype=np.float64)
    my, mz = a.shape
    b = np.zeros((my // 2 + 1, mz), dtype=np.complex128, order="F")
    load().orc_ff2rc(my, mz, _p(a), _p(b))
    return b


def ff2cr(b, my):
    b = np.asfortranarray(b, dtype=np.complex128)
    mz = b.shape[1]
    a = np.zeros((my, mz), order="F")
    load().orc_ff2cr(my, mz, _p(b), _p(a))
    return a


def fft1d(x, sign=-1):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    load().orc_fft1d(len(x), sign, _p(x), _p(out))
    return out


def tridag(a, b, c, r):
    a, b, c, r = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, r))
    u = np.zeros_like(r)
    st = load().orc_tridag(_p(a), _p(b), _p(c), _p(r), _p(u), len(r))
    return u, st


def cyclic_solve(a, b, c, r):
    a, b, c, r = (np.ascontiguousarray(v, dtype=np.float64) for v in (a, b, c, r))
    x = np.zeros_like(r)
    st = load().orc_cyclic_solve(_p(a), _p(b), _p(c), _p(r), _p(x), len(r))
    return x, st


def quadgener(r, f2=1.6, dr=1.0, f1=1.6):
    n = load().orc_quadgener(r, dr, f1, f2, 0, None, None)
    bnd = np.zeros((3, n), order="F")
    cd = np.zeros((3, n), order="F")
    load().orc_quadgener(r, dr, f1, f2, n, _p(bnd), _p(cd))
    return bnd, cd


def if_rev(solver):
    nb, nbnd = solver.setup.nbody, solver.setup.nbnd
    out = np.zeros((nbnd, nb), dtype=np.uint8, order="F")
    load().orc_get_if_rev(solver.ctx, out.ctypes.data_as(ctypes.POINTER(ctypes.c_ubyte)))
    return out


def dem_neighbours(dem):
    """NEIGHBOURS(l, 1..) of an oracle DEM context as a list of sorted sets of
    0-based partner indices (the particle itself left out)."""
    n = dem.nbody
    buf = np.zeros(n * 64, dtype=np.int32)
    w = load().orc_dem_neighbours(dem.ctx, buf.ctypes.data_as(ctypes.POINTER(ctypes.c_int)))
    out = []
    for l in range(n):
        row = buf[l * w:(l + 1) * w]
        cnt = row[0]
        out.append(sorted(int(v) - 1 for v in row[1:cnt + 1] if v - 1 != l))
    return out
