"""The Fortran driver (fsrc/pib_main.f90) end to end: a namelist deck in,
the reference's _norm_drag.dat / _drag_components.dat records out, compared
with the oracle advanced over the same steps (norm_drag to 1e-9)."""
import ctypes
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


def test_driver_moves_an_imove_bed(tmp_path, orc):
    """IMOVE = 1: the driver runs the moving-bed stages (drag, GPU DEM, new
    mask and reversal pairs, forcing, velstep); its drag records equal the
    oracle chain's (oracle fluid, oracle DEM) to 1e-9, and the particles'
    mean velocity column is no longer zero."""
    print("fortran driver, moving 32^3 bed, 4 steps", flush=True)
    kw = dict(kind="random", my=32, dbydx=8, nbody=6, seed=3, imove=1, rhos=2.0, coeff_rest=0.8)
    st = make_setup(tmp_path, "DMV", **kw)
    nsteps = 4
    run = os.path.join(str(tmp_path), "DMV")
    r = subprocess.run([DRIVER, run, str(nsteps), "fp64"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0 and "MOVING PARTICLES" in r.stdout
    nd = np.loadtxt(run + "_norm_drag.dat", ndmin=2)
    assert nd.shape[0] == nsteps
    st._load()
    host = _abi.load_host()
    dc = _abi.PibDemConfig()
    assert host.pib_host_get_dem_config(ctypes.byref(dc)) == 0
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    do = solver.Dem(st.xc, st.radbdy, dx=dc.dx, xlength=dc.xlength, ylength=dc.ylength, zlength=dc.zlength,
                    foffset=dc.foffset, rhos=dc.rhos, rhof=dc.rhof, dia_phys=dc.dia_phys,
                    coeff_rest=dc.coeff_rest, lib=orc.load(), prefix="orc_")
    hist = []
    for n in range(nsteps):
        o.step_moving(do, 1)
        hist.append(o.scalars().norm_drag)
    hist = np.array(hist)
    assert np.max(np.abs(nd[:, 3] - hist) / np.abs(hist)) < 1e-9
    assert np.any(np.abs(nd[-1, 5:8]) > 0)       # usmean_act of the moving particles


def test_driver_runs_a_scalar_deck(tmp_path, orc):
    """ISCALON = 1: each stage ends with the scalar step; the driver's
    _scalflux.dat and _nusselt_comps.dat records equal the oracle chain's
    (same stage calls, tscal advanced by dt each stage) to 1e-9."""
    print("fortran driver, 24^3 bed with the scalar, 5 steps", flush=True)
    kw = dict(kind="random", my=24, dbydx=6, nbody=3, seed=9, re=20.0, discr_scheme="upwind", iscalon=1,
              phisurf=0.0, phistream=1.0, pr=0.7)
    st = make_setup(tmp_path, "DSC", **kw)
    nsteps = 5
    run = os.path.join(str(tmp_path), "DSC")
    r = subprocess.run([DRIVER, run, str(nsteps), "fp64"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0 and "SCALAR ON" in r.stdout
    flux = np.loadtxt(run + "_scalflux.dat", ndmin=2)
    nuss = np.loadtxt(run + "_nusselt_comps.dat", ndmin=2)
    assert flux.shape == (nsteps, 3 + st.nbody + 1) and nuss.shape == (nsteps, 5)
    host = _abi.load_host()
    tc, tv, td = (ctypes.c_double() for _ in range(3))
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    host.pib_host_time_scales(ctypes.byref(tc), ctypes.byref(tv), ctypes.byref(td))
    sc = st.scalar
    dphi = sc["phisurf"] - sc["phistream"]
    tscal = 0.0
    for n in range(nsteps):
        o.compute_new_timestep(1)
        tscal += o.scalars().dt
        o.nonlinear(1)
        o.velstep(1)
        o.scalstep(1)
        fb, fg, _, _ = o.scalar_flux()
        nus = -(fg * sc["char_length"] ** 2) / (6 * sc["voldom"] * o.scalars().maxvolfrac * dphi)
        ref = np.concatenate([[tscal / tc.value, tscal / tv.value, tscal / td.value], -fb / dphi, [-fg / dphi]])
        assert np.max(np.abs(flux[n] - ref) / np.maximum(np.abs(ref), 1e-30)) < 1e-9, n
        assert abs(nuss[n, 3] - nus) <= 1e-9 * abs(nus)
        assert abs(nuss[n, 4] - o.scalar_get()[2]) <= 1e-11 * abs(o.scalar_get()[2])  # e20.12 text
    print(f"  Nusselt after {nsteps} steps: {nuss[-1, 3]:.10e}, heat ratio {nuss[-1, 4]:.12f}", flush=True)
    o.close()


def test_driver_stops_when_the_drag_error_history_converges(tmp_path, orc):
    """dtibm_main_scal_flo.F90:471-476, 652-658: a fixed bed stops once the
    mean of the NERR_STEPS drag-error history is below TOL_FERROR, after
    writing that step's records and RUN_NAME_CONVERGED.  NERR_STEPS = 2 and
    TOL_FERROR = 0.99 make the stop fall inside the run; the step it falls on
    is the first one whose oracle history mean is below the tolerance."""
    print("fortran driver, 24^3 bed, convergence stop", flush=True)
    assert os.path.exists(DRIVER), "driver not built"
    kw = dict(kind="random", my=24, dbydx=6, nbody=3, seed=5, re=40.0, nerr_steps=2, tol_ferror=0.99)
    st = make_setup(tmp_path, "DCV", **kw)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    o.set_ferror_history(2)
    nsteps, stop = 8, None
    for n in range(nsteps):
        o.step(1)
        if o.ferror_history(2)[1] < 0.99:
            stop = n + 1
            break
    assert stop is not None and stop < nsteps
    run = os.path.join(str(tmp_path), "DCV")
    r = subprocess.run([DRIVER, run, str(nsteps), "fp64"], capture_output=True, text=True, timeout=120)
    print(r.stdout[-1500:], r.stderr[-1500:], flush=True)
    assert r.returncode == 0
    nd = np.loadtxt(run + "_norm_drag.dat", ndmin=2)
    assert nd.shape[0] == stop
    assert os.path.exists(run + "_CONVERGED")
    assert "FERROR_HIST AND TOL" in r.stdout
