"""The Fortran driver (fsrc/pib_main.f90) end to end: a namelist deck in,
the reference's _norm_drag.dat / _drag_components.dat records out, compared
with the oracle advanced over the same steps (norm_drag to 1e-9)."""
import os
import subprocess

import numpy as np
import pytest

from conftest import make_setup
from pureibm_ps2d_fd1d_amd import _abi, solver

pytestmark = pytest.mark.gpu

DRIVER = os.path.join(os.path.dirname(_abi.gpu_lib_path()), "pureibm_gpu")


def test_driver_writes_reference_drag_records(tmp_path, orc):
    print("fortran driver, 24^3 bed, 5 steps", flush=True)
    assert os.path.exists(DRIVER), "driver not built"
    kw = dict(kind="random", my=24, dbydx=6, nbody=3, seed=5, re=40.0)
    st = make_setup(tmp_path, "DRV", **kw)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    nsteps = 5
    # the same single-step calls from Python, for comparison with the driver
    g = solver.Solver(st)
    for n in range(nsteps):
        g.step(1)
        sg = g.scalars()
        print(f"  python step {n + 1}: norm_drag {sg.norm_drag:.10e} t {sg.t:.6e} mpg {sg.mpg[0]:.6e} "
              f"ufmean {sg.ufmean[0]:.6e} ufmean_des {sg.ufmean_des[0]:.6e}", flush=True)
    g.close()
    run = os.path.join(str(tmp_path), "DRV")
    r = subprocess.run([DRIVER, run, str(nsteps), "fp64"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-2000:], r.stderr[-2000:], flush=True)
    assert r.returncode == 0
    nd = np.loadtxt(run + "_norm_drag.dat", ndmin=2)
    comps = np.loadtxt(run + "_drag_components.dat", ndmin=2)
    assert nd.shape[0] == nsteps and comps.shape[0] == nsteps
    hist = []
    for n in range(nsteps):
        o.step(1)
        hist.append(o.scalars().norm_drag)
    hist = np.array(hist)
    assert np.max(np.abs(nd[:, 3] - hist) / np.abs(hist)) < 1e-9
    # drag components: norm_drag in column 5, pres + visc = total in columns 1..3
    assert np.allclose(comps[:, 4], nd[:, 3], rtol=1e-12)
    assert np.allclose(comps[:, 0] + comps[:, 1], comps[:, 2], rtol=1e-12)
    # columns: t/t_conv/(1-phi), t/t_vis, t/t_diff, norm_drag, norm_drag (one phase),
    # usmean_act(1:3), ufmean(1:3)/ucharmod
    so = o.scalars()
    uf = so.ufmean[0] / st.config.ucharmod
    assert abs(nd[-1, 8] - uf) <= 1e-9 * abs(uf)
    assert nd.shape[1] == 11
