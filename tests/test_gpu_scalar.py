"""Scalar transport, one species: pib_scalstep_modes (the per-mode solve of
scalstep, stepthescalar.F90:67-241) against the oracle on random spectral
state, over shapes with ragged mode counts, odd nx and long lines, both
context precisions (the scalar solve is FP64 in both), AB2 and RK3 stages,
and heat_ratio != heat_ratio_new with the ratios rolled across two calls.
On the 128^3 and 256^3 grids the Crank-Nicolson factor of cos modes is
checked in closed form.

Tolerance: relative 1e-12 to the field maximum (both sides FP64 Thomas +
Sherman-Morrison; the GPU multiplies by one reciprocal pivot a row where the
reference divides each right-hand side, and fuses multiply-adds, so the two
round differently).
"""
import ctypes
import math

import numpy as np
import pytest

from conftest import rel_err
from pureibm_ps2d_fd1d_amd import _abi
from test_gpu_velstep import Bare, _cfg

pytestmark = pytest.mark.gpu


def _ratios(b):
    out = (ctypes.c_double * 4)()
    assert b.fn("scalar_get")(b.ctx, out) == 0
    return tuple(out)


def _set_dt(b, dt):
    s = _abi.PibScalars()
    assert b.fn("get_scalars")(b.ctx, ctypes.byref(s)) == 0
    s.dt = dt
    assert b.fn("set_scalars")(b.ctx, ctypes.byref(s)) == 0


@pytest.mark.parametrize("prec", [0, 1])
@pytest.mark.parametrize("nx,my,mz,rk3", [
    (8, 8, 8, 0), (12, 8, 6, 1), (33, 12, 10, 0), (16, 16, 16, 1), (128, 16, 16, 0), (256, 40, 24, 1),
])
def test_scalstep_random_state(orc, nx, my, mz, rk3, prec):
    print(f"scalstep nx={nx} my={my} mz={mz} rk3={rk3} prec={prec}", flush=True)
    cfg = _cfg(nx, my, mz, prec, rk3)
    o = Bare(orc.load(), "orc_", cfg)
    g = Bare(_abi.load_gpu(), "pib_", cfg)
    my2 = my // 2 + 1
    shp = (nx + 1, my2, mz)
    rng = np.random.default_rng(nx * 131 + my * 7 + mz)
    gamma, hr, hrn = 0.021, 1.07, 0.93
    for b in (o, g):
        assert b.fn("scalar_set")(b.ctx, gamma, hr, hrn) == 0
        _set_dt(b, 0.0137)
    for call, rks in enumerate((1, 2) if rk3 else (1, 1)):
        for f in ("phi", "nlphi", "onlphi", "ffphi"):
            if f == "phi" and call:
                continue                         # the second call starts from the first's result
            a = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
            o.set(f, a)
            g.set(f, a)
        assert o.fn("scalstep_modes")(o.ctx, rks) == 0
        assert g.fn("scalstep_modes")(g.ctx, rks) == 0
        po = o.get("phi", shp, np.complex128)
        pg = g.get("phi", shp, np.complex128)
        e = rel_err(pg, po)
        print(f"  call {call} rks {rks}: phi {e:.2e}", flush=True)
        assert e < 1e-12
        assert _ratios(g) == _ratios(o)
        if call == 0:
            o.fn("scalar_set")(o.ctx, gamma, hrn, 1.0 / hrn)
            g.fn("scalar_set")(g.ctx, gamma, hrn, 1.0 / hrn)
    o.close()
    g.close()


@pytest.mark.parametrize("n", [128, 256])
def test_scalstep_cos_modes_full_size(n):
    print(f"scalstep cos modes {n}^3", flush=True)
    nx = my = mz = n
    cfg = _cfg(nx, my, mz, 0, 0)
    g = Bare(_abi.load_gpu(), "pib_", cfg)
    my2 = my // 2 + 1
    shp = (nx + 1, my2, mz)
    gamma, dt, dx = 0.013, 0.004, cfg.dx
    assert g.fn("scalar_set")(g.ctx, gamma, 1.0, 1.0) == 0
    _set_dt(g, dt)
    rng = np.random.default_rng(11)
    m = rng.integers(0, nx // 2 + 1, size=shp[1:])
    amp = rng.standard_normal(shp[1:]) + 1j * rng.standard_normal(shp[1:])
    i = np.arange(nx + 1)[:, None, None]
    phi = amp * np.cos(2 * math.pi * m * i / nx + 0.3)
    g.set("phi", phi)
    assert g.fn("scalstep_modes")(g.ctx, 1) == 0
    got = g.get("phi", shp, np.complex128)
    ky = 2 * math.pi * np.arange(my2) / (my * dx)
    kz = 2 * math.pi * np.fft.fftfreq(mz, d=1.0 / mz) / (mz * dx)
    lam = 4 * np.sin(math.pi * m / nx) ** 2 / dx ** 2 + ky[:, None] ** 2 + kz[None, :] ** 2
    fac = (1 - dt * 0.5 * gamma * lam) / (1 + dt * 0.5 * gamma * lam)
    assert rel_err(got, fac * phi) < 1e-11
    g.close()


def test_scalar_state_errors():
    g = Bare(_abi.load_gpu(), "pib_", _cfg(8, 8, 8, 0, 0))
    assert g.fn("scalstep_modes")(g.ctx, 1) == _abi.PIB_ERR_STATE
    a = np.zeros((9, 5, 8), dtype=np.complex128, order="F")
    ptr = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert g.fn("get_field")(g.ctx, _abi.FIELDS["phi"], ptr) == _abi.PIB_ERR_STATE
    assert g.fn("set_field")(g.ctx, _abi.FIELDS["nlphi"], ptr) == _abi.PIB_ERR_STATE
    assert g.fn("scalar_get")(g.ctx, (ctypes.c_double * 4)()) == _abi.PIB_ERR_STATE
    assert g.fn("scalar_set")(g.ctx, -1.0, 1.0, 1.0) == _abi.PIB_ERR_ARG
    assert g.fn("scalar_set")(g.ctx, 0.5, 1.0, 1.0) == 0
    assert g.fn("scalstep_modes")(g.ctx, 2) == _abi.PIB_ERR_ARG
    assert g.fn("scalstep_modes")(g.ctx, 1) == 0
    g.close()


def test_scalstep_after_fluid_steps_uses_the_steps_dt(tmp_path, orc):
    """Five fluid steps of a random bed (adaptive dt) in two pib_step calls --
    steps 3 and 4 replay the captured step graph, step 5 (the last of a call)
    runs eagerly from the state the replays wrote -- then the scalar solve: it
    reads dt from the context's device state, as the reference's phistep reads
    the module's dt, and agrees with the oracle run through the same calls.
    The replayed path is the one pib_step takes from a context's third step on
    (graph replays are bitwise equal to eager steps, tests/test_gpu_graph.py)."""
    from conftest import make_setup
    from pureibm_ps2d_fd1d_amd import solver
    print("scalstep after 5 fluid steps (2 replayed from a graph), 24^3 bed", flush=True)
    st = make_setup(tmp_path, "SC", kind="random", my=24, dbydx=6, nbody=3, seed=4, re=30.0)
    g = solver.Solver(st)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    g.step(2)
    g.step(3)
    o.step(5)
    assert abs(g.scalars().dt - o.scalars().dt) <= 1e-12 * o.scalars().dt
    shp = (st.nx + 1, st.my // 2 + 1, st.mz)
    rng = np.random.default_rng(9)
    for b in (g, o):
        b.scalar_set(0.05, 1.02, 0.98)
    for f in ("phi", "nlphi", "onlphi", "ffphi"):
        a = rng.standard_normal(shp) + 1j * rng.standard_normal(shp)
        g.set_field(f, a)
        o.set_field(f, a)
    g.scalstep_modes(1)
    o.scalstep_modes(1)
    assert g._fn("sync")(g.ctx) == 0  # the solve's device status
    e = rel_err(g.field("phi"), o.field("phi"))
    print(f"  phi {e:.2e}", flush=True)
    assert e < 1e-11
    assert g.scalar_get() == o.scalar_get()
    g.close()
    o.close()
