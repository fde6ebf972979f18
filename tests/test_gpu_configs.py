"""GPU parity at the benchmark workloads themselves, through the C-ABI.

* C1: 64^3, one sphere of 16 cells (simple cubic), Re_m = 20, fixed dt,
  200 steps against the oracle, compared every 50 steps.
* C2: 128^3, 195 RSA spheres of 16 cells (phi = 0.20), Re_m = 50: FP64
  against the oracle for 2 steps, and FP32-bulk against FP64 over the full
  500 steps (mean drag within 1e-4).
* C3: a 96^3 and a 160^3 realization of the batched workload at their real
  sphere counts (upwind x-fluxes) against the oracle for 2 steps.
* Stokes flow past a dilute simple cubic array: the normalised drag tends to
  the Hasimoto / Sangani-Acrivos value as the sphere is resolved, a check of
  the drag normalisation and the mean-pressure-gradient controller against
  the physics rather than against the oracle.

Tolerances (FP64 against the oracle): relative 1e-10 on fields, 1e-9 on
per-particle drag.
"""
import math
import os

import numpy as np
import pytest

from conftest import rel_err
from pureibm_ps2d_fd1d_amd import _abi, configs, solver

pytestmark = pytest.mark.gpu


def _setup(tmp_path, name, precision=_abi.PIB_FP64, **over):
    d = str(tmp_path)
    configs.write_workload(d, name, **over)
    return solver.RunSetup(os.path.join(d, name), precision=precision)


def _compare(g, o, tag):
    eu = rel_err(g.field("u"), o.field("u"))
    ep = rel_err(g.field("p"), o.field("p"))
    ef = rel_err(g.forces()["force"], o.forces()["force"])
    sg, so = g.scalars(), o.scalars()
    en = abs(sg.norm_drag - so.norm_drag) / abs(so.norm_drag)
    print(f"  {tag}: u {eu:.2e} p {ep:.2e} force {ef:.2e} norm_drag {so.norm_drag:.10f} ({en:.1e})", flush=True)
    assert eu < 1e-10 and ep < 1e-10, tag
    assert ef < 1e-9 and en < 1e-9, tag
    assert abs(sg.t - so.t) <= 1e-12 * abs(so.t)
    for n in range(3):
        assert abs(sg.mpg[n] - so.mpg[n]) <= 1e-9 * max(abs(so.mpg[0]), 1e-300)


def test_c1_200_steps_against_the_oracle(tmp_path, orc):
    """The C1 workload: a simple cubic bed of one sphere, ADAPTIVE = F, RAMP_FRAC_STEPS = 0."""
    print("C1 64^3 one sphere, 200 steps", flush=True)
    st = _setup(tmp_path, "C1", adaptive=False, ramp_frac_steps=0)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    g = solver.Solver(st)
    dt0 = g.scalars().dt
    for k in range(4):
        o.step(50)
        g.step(50)
        _compare(g, o, f"step {50 * (k + 1)}")
    assert g.scalars().dt == dt0  # fixed time step
    g.close()
    o.close()


def test_c2_fp64_against_the_oracle(tmp_path, orc):
    print("C2 128^3 195 spheres, FP64, 2 steps", flush=True)
    st = _setup(tmp_path, "C2")
    assert st.nbody == 195
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    g = solver.Solver(st)
    assert g.count_solid == o.count_solid
    assert np.array_equal(g.nrpr_active, o.nrpr_active)
    for k in range(2):
        o.step(1)
        g.step(1)
        _compare(g, o, f"step {k + 1}")
    g.close()
    o.close()


def test_c2_fp32_bulk_500_steps_against_fp64(tmp_path):
    """The north-star gate: FP32-bulk against FP64 on the C2 bed after the
    full 500 steps, relative mean-drag error within 1e-4."""
    print("C2 FP32-bulk vs FP64, 500 steps", flush=True)
    g64 = solver.Solver(_setup(tmp_path / "a", "C2"))
    g32 = solver.Solver(_setup(tmp_path / "b", "C2", precision=_abi.PIB_FP32_BULK))
    for k in range(5):
        g64.step(100)
        g32.step(100)
        s64, s32 = g64.scalars(), g32.scalars()
        f64, f32 = g64.forces()["force"], g32.forces()["force"]
        mean_err = abs(s32.norm_drag - s64.norm_drag) / abs(s64.norm_drag)
        fx_err = abs(f32[:, 0].mean() - f64[:, 0].mean()) / abs(f64[:, 0].mean())
        part_err = np.max(np.abs(f32 - f64)) / np.max(np.abs(f64))
        print(f"  step {100 * (k + 1)}: norm_drag {s64.norm_drag:.8f} / {s32.norm_drag:.8f}, mean drag "
              f"rel err {mean_err:.2e} (x: {fx_err:.2e}), largest per-particle {part_err:.2e}", flush=True)
    assert mean_err <= 1e-4 and fx_err <= 1e-4
    assert part_err <= 1e-3
    g64.close()
    g32.close()


@pytest.mark.parametrize("k", [0, 2])
def test_c3_realization_against_the_oracle(tmp_path, orc, k):
    base = configs.c3_realization(str(tmp_path), k)
    st = solver.RunSetup(base)
    print(f"C3 realization {k}: {st.nx}^3, {st.nbody} spheres, Re_m from the deck, 2 steps", flush=True)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    g = solver.Solver(st)
    assert g.count_solid == o.count_solid
    for q in range(2):
        o.step(1)
        g.step(1)
        _compare(g, o, f"step {q + 1}")
    g.close()
    o.close()


def _sangani_acrivos(phi):
    """Dimensionless Stokes drag F/(3 pi mu d U) of a simple cubic array
    (Hasimoto 1959, extended by Sangani & Acrivos 1982)."""
    c = phi ** (1.0 / 3.0)
    return 1.0 / (1.0 - 1.7601 * c + phi - 1.5593 * phi ** 2 + 3.9799 * c ** 8 - 3.0734 * c ** 10)


def test_stokes_drag_of_a_simple_cubic_array(tmp_path):
    """One sphere in a periodic box (phi = 0.0082) at Re_m = 0.1, run to a
    steady state with small steps: the normalised drag norm_drag
    = |F| / (3 pi mu U d) tends to Sangani & Acrivos' value as the sphere is
    resolved (8 and 16 cells a diameter), so a wrong factor in the drag
    normalisation, the surface quadrature or the mpg controller shows."""
    res = {}
    for my, dbydx, nsteps in ((32, 8, 1500), (64, 16, 2400)):
        d = str(tmp_path / f"s{dbydx}")
        configs.simple_cubic(d, "ST", my=my, dbydx=dbydx, re=0.1, cfl=0.002, cfl_max=0.002)
        st = solver.RunSetup(os.path.join(d, "ST"))
        g = solver.Solver(st)
        hist = []
        for q in range(nsteps // 100):
            g.step(100)
            hist.append(g.scalars().norm_drag)
        phi = math.pi / 6.0 * (dbydx / my) ** 3
        K = _sangani_acrivos(phi)
        res[dbydx] = (hist[-1], K)
        print(f"  d = {dbydx} cells: norm_drag {hist[-1]:.5f} (last 100 steps changed it by "
              f"{abs(hist[-1] - hist[-2]) / hist[-1]:.1e}), Sangani-Acrivos {K:.5f}, "
              f"rel diff {(hist[-1] - K) / K:+.3f}", flush=True)
        assert abs(hist[-1] - hist[-2]) < 2e-3 * hist[-1]  # steady
        g.close()
    e8 = abs(res[8][0] - res[8][1]) / res[8][1]
    e16 = abs(res[16][0] - res[16][1]) / res[16][1]
    assert e8 < 0.08 and e16 < 0.05
    assert e16 < e8
