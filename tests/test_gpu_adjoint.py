"""The hand-written adjoint of the fixed-bed step on the GPU, through the C-ABI.

* One adjoint step from the same state and the same adjoint seeds equals the
  oracle's (test_adjoint_oracle.py pins the oracle's) on generic and compiled
  plane shapes, FP64: relative 1e-9 (the GPU sums the transposed
  interpolations with atomics, in no fixed order).
* The GPU adjoint step is the transpose of the GPU forward step: a central
  difference of the quadratic step along a random direction, FP64, 1e-9.
* The adjoint workload C4 (128^3, 50 spheres of 16 cells, Re_m = 20,
  SET_MPG, fixed dt, 64 steps checkpointed every 8): dJ/dxc for all 150
  centre coordinates and dJ/dmpg of J = total streamwise drag after the 64th
  step against central finite differences with h = 1e-6 cell in FP64,
  within 1e-5 of the largest gradient entry.  Coordinates whose surface,
  inner, outer or pressure points change stencil cell, or whose mask or
  active reversal pairs change, within +-h are reported and skipped.
* The same gradient in FP32-bulk, and through torch.autograd.
"""
import os
import time

import numpy as np
import pytest

from conftest import rel_err
from test_adjoint_oracle import _setup, fd_gradient, transpose_check
from pureibm_ps2d_fd1d_amd import _abi, adjoint, configs, solver

pytestmark = pytest.mark.gpu


def _seeded_pair(st, orc, rng, nsteps_before=3):
    g = solver.Solver(st)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    sc = o.scalars()
    for n in range(3):
        sc.umean[n] = sc.ufmean_des[n]
    o.set_scalars(sc)
    o.step(nsteps_before)
    for f in ("u", "p", "nl"):
        g.set_field(f, o.field(f))
    g.set_scalars(o.scalars())
    shp = {f: o.field(f).shape for f in ("u", "p", "nl")}
    seeds = {f: rng.standard_normal(shp[f[1:]]) + 1j * rng.standard_normal(shp[f[1:]]) for f in ("au", "ap", "anl")}
    yum = rng.standard_normal(3)
    wf = np.asfortranarray(rng.standard_normal((st.nbody, 3)))
    for b in (g, o):
        b.adjoint_begin()
        for f, a in seeds.items():
            b.set_field(f, a)
        b.adjoint_set_umean(yum)
    return g, o, wf


@pytest.mark.parametrize("my,dbydx,nbody", [(16, 6, 2), (24, 6, 4), (32, 8, 3)])
def test_adjoint_step_matches_the_oracle(tmp_path, orc, my, dbydx, nbody):
    st = _setup(tmp_path, my=my, dbydx=dbydx, nbody=nbody, seed=11)
    g, o, wf = _seeded_pair(st, orc, np.random.default_rng(3))
    print(f"adjoint step {my}^3, {nbody} spheres", flush=True)
    for b in (g, o):
        b.adjoint_step(wf)
        b.adjoint_step(None)
    nx = g.nx
    for f in ("au", "ap", "anl"):
        e = rel_err(g.field(f)[:nx], o.field(f)[:nx])
        print(f"  {f}: {e:.2e}", flush=True)
        assert e < 1e-9, f
    (gx, gm, gu), (ox, om, ou) = g.adjoint_get(), o.adjoint_get()
    print(f"  dxc {rel_err(gx, ox):.2e} dmpg {rel_err(gm, om):.2e} dumean {rel_err(gu, ou):.2e}", flush=True)
    assert rel_err(gx, ox) < 1e-9 and rel_err(gm, om) < 1e-9 and rel_err(gu, ou) < 1e-9
    g.close()
    o.close()


def test_gpu_adjoint_step_is_the_transpose_of_the_gpu_step(tmp_path):
    st = _setup(tmp_path, my=32, dbydx=8, nbody=3, seed=5)
    s = solver.Solver(st)
    fd, ad = transpose_check(s, np.random.default_rng(9), s.nx)
    print(f"central difference {fd:.15e}, adjoint {ad:.15e}", flush=True)
    assert abs(fd - ad) <= 1e-9 * abs(fd)
    s.close()


@pytest.mark.parametrize("prec", [_abi.PIB_FP64, _abi.PIB_FP32_BULK])
def test_backward_sweep_on_recorded_steps_matches_the_oracle(tmp_path, orc, prec):
    """The whole backward sweep (checkpoint pops, pib_adjoint_record, adjoint
    steps on recorded fields and on recomputed ones) against the oracle's,
    with force weights on steps inside segments as well as at their ends, so
    records that keep their own onlbc sit above the bottom of the stack.
    Then, on fresh contexts, steps recorded between plain steps: the state
    after them must be that of plain steps (the context's buffers hold the
    last recorded step's fields).  FP64: relative 1e-9; FP32-bulk (GPU
    only, the oracle is FP64): 1e-4."""
    st = _setup(tmp_path, my=24, dbydx=6, nbody=4, seed=7)
    nsteps, every = 14, 5
    rng = np.random.default_rng(5)
    w = {n: np.asfortranarray(rng.standard_normal((st.nbody, 3))) for n in (3, 9, 13)}
    out = []
    for lib in ("gpu", "orc"):
        st.config.precision = prec if lib == "gpu" else _abi.PIB_FP64

        def fresh():
            s = solver.Solver(st) if lib == "gpu" else solver.Solver(st, lib=orc.load(), prefix="orc_")
            sc = s.scalars()
            for n in range(3):
                sc.umean[n] = sc.ufmean_des[n]
            s.set_scalars(sc)
            return s
        s = fresh()
        adjoint.forward_checkpointed(s, nsteps, every)
        grad = adjoint.backward(s, nsteps, every, w.get)
        s.close()
        s = fresh()
        s.step(3)
        s.adjoint_begin()
        s.adjoint_record(3)
        s.step(1)
        s.adjoint_record(4)
        s.step(2)
        out.append((grad, s.field("u"), s.field("p"), s.field("nl")))
        s.close()
    (gg, gu, gp, gn), (og, ou, op, on) = out
    tol = 1e-9 if prec == _abi.PIB_FP64 else 1e-4
    errs = [rel_err(a, b) for a, b in zip(gg, og)] + [rel_err(gu, ou), rel_err(gp, op), rel_err(gn, on)]
    print("backward sweep: dxc {:.1e} dmpg {:.1e} dumean {:.1e}; recorded then plain steps: u {:.1e} p {:.1e} "
          "nl {:.1e}".format(*errs), flush=True)
    assert max(errs) < tol


def test_gpu_gradient_matches_finite_differences_small_bed(tmp_path):
    st = _setup(tmp_path, my=24, dbydx=6, nbody=3, seed=2)
    P = adjoint.DragProblem(st, nsteps=6, every=4)
    xc0, mpg0 = st.xc.copy(), np.array([0.3, 0.05, -0.02])
    P.objective(xc0, mpg0)
    dxc, dmpg = P.gradient()
    fd, fdm, skip = fd_gradient(P, st, xc0, mpg0, 1e-6)
    ok = ~np.isnan(fd)
    print(f"checked {ok.sum()} coordinates, skipped {skip}", flush=True)
    assert ok.sum() >= 6
    assert np.max(np.abs(dxc[ok] - fd[ok])) <= 1e-6 * np.max(np.abs(fd[ok]))
    assert np.max(np.abs(dmpg - fdm)) <= 1e-6 * np.max(np.abs(fdm))
    P.close()


def _c4(tmp_path, precision=_abi.PIB_FP64):
    d = str(tmp_path)
    configs.write_workload(d, "C4")
    return solver.RunSetup(os.path.join(d, "C4"), precision=precision)


def test_c4_gradient_against_finite_differences(tmp_path):
    st = _c4(tmp_path)
    P = adjoint.DragProblem(st, nsteps=64, every=8)
    xc0 = st.xc.copy()
    P.objective(xc0, None, keep=False)
    mpg0 = adjoint.stokes_mpg(P.s)
    t0 = time.time()
    J = P.objective(xc0, mpg0)
    dxc, dmpg = P.gradient()
    print(f"C4: J = {J:.10e}, gradient in {time.time() - t0:.2f} s", flush=True)
    t0 = time.time()
    fd, fdm, skip = fd_gradient(P, st, xc0, mpg0, 1e-6)
    ok = ~np.isnan(fd)
    scale = np.max(np.abs(fd[ok]))
    err = np.max(np.abs(dxc[ok] - fd[ok])) / scale
    errm = np.max(np.abs(dmpg - fdm)) / np.max(np.abs(fdm))
    print(f"C4: {ok.sum()} of 150 coordinates checked ({len(skip)} skipped: {skip}) in {time.time() - t0:.1f} s; "
          f"max |adjoint - fd| / max|fd| = {err:.2e}; mpg {errm:.2e}", flush=True)
    assert ok.sum() >= 140
    assert err <= 1e-5 and errm <= 1e-5
    np.save(os.path.join(str(tmp_path), "c4_grad.npy"), dxc)
    P.close()


def test_c4_fp32_bulk_gradient_against_fp64(tmp_path):
    g = {}
    for prec in (_abi.PIB_FP64, _abi.PIB_FP32_BULK):
        st = _c4(tmp_path, prec)
        P = adjoint.DragProblem(st, nsteps=64, every=8)
        P.objective(None, None, keep=False)
        mpg0 = adjoint.stokes_mpg(P.s)
        P.objective(None, mpg0)
        g[prec] = P.gradient()
        P.close()
    e = rel_err(g[_abi.PIB_FP32_BULK][0], g[_abi.PIB_FP64][0])
    em = rel_err(g[_abi.PIB_FP32_BULK][1], g[_abi.PIB_FP64][1])
    print(f"C4 FP32-bulk gradient vs FP64: dxc {e:.2e}, dmpg {em:.2e}", flush=True)
    assert e < 1e-2 and em < 1e-2


def test_gpu_torch_function(tmp_path):
    torch = pytest.importorskip("torch")
    st = _setup(tmp_path, my=24, dbydx=6, nbody=3, seed=2)
    P = adjoint.DragProblem(st, nsteps=4, every=2)
    xc = torch.tensor(st.xc, dtype=torch.float64, requires_grad=True)
    mpg = torch.tensor([0.3, 0.05, -0.02], dtype=torch.float64, requires_grad=True)
    J = adjoint.drag_objective().apply(xc, mpg, P)
    J.backward()
    P2 = adjoint.DragProblem(st, nsteps=4, every=2)
    J2 = P2.objective(st.xc, np.array([0.3, 0.05, -0.02]))
    dxc, dmpg = P2.gradient()
    assert float(J.detach()) == J2
    assert rel_err(xc.grad.numpy(), dxc) < 1e-12 and rel_err(mpg.grad.numpy(), dmpg) < 1e-12
    P.close()
    P2.close()
