"""Batched realizations through the C-ABI.

* Independent contexts launched together (pib_step_async on each, then
  pib_sync) give bitwise the same state as the same contexts stepped one
  after the other: the contexts share nothing on the device.
* pib_batch_step (one graph launch a step over every context's step) gives
  bitwise the state of each context stepped alone, for contexts of several
  sizes and both precisions, across a change of one context's particles
  (its graphs dropped, the batch graph rebuilt).
* Two realizations of the batched workload C3 (96^3, 216 and 109 spheres,
  upwind x-fluxes), FP64, stepped by pib_batch_step against the oracle for 4
  steps (the last two are batch-graph launches): relative 1e-10 on fields,
  1e-9 on per-particle drag.
* pib_autotune leaves the state as it found it, picks forms within range,
  caches them per shape, and the tuned context steps to the oracle's result.
"""
import numpy as np
import pytest

from conftest import make_setup, rel_err
from pureibm_ps2d_fd1d_amd import _abi, configs, solver

pytestmark = pytest.mark.gpu


def test_concurrent_contexts_equal_sequential(tmp_path):
    print("concurrent contexts vs sequential", flush=True)
    cases = [dict(kind="random", my=24, dbydx=6, nbody=3, seed=21, re=30.0),
             dict(kind="random", my=32, dbydx=8, nbody=4, seed=22, re=40.0),
             dict(kind="random", my=20, dbydx=5, nbody=2, seed=23, re=20.0)]
    sts = [make_setup(tmp_path / f"c{q}", f"B{q}", precision=_abi.PIB_FP32_BULK, **kw) for q, kw in enumerate(cases)]
    seq = [solver.Solver(st) for st in sts]
    par = [solver.Solver(st) for st in sts]
    for g in seq:
        g.step(6)
    for g in par:
        g.step(2)
    for g in par:
        assert g.lib.pib_step_async(g.ctx, 4) == 0
    for g in par:
        assert g.lib.pib_sync(g.ctx) == 0
    for a, b in zip(seq, par):
        for f in ("u", "p", "nl"):
            assert np.array_equal(a.field(f), b.field(f)), f
        assert a.scalars().norm_drag == b.scalars().norm_drag
    for g in seq + par:
        g.close()


def _same(a, b, tag):
    for f in ("u", "p", "nl"):
        assert np.array_equal(a.field(f), b.field(f)), (tag, f)
    assert a.scalars().norm_drag == b.scalars().norm_drag, tag
    assert np.array_equal(a.forces()["force"], b.forces()["force"]), tag


def test_batch_graph_equals_contexts_stepped_alone(tmp_path):
    cases = [dict(kind="random", my=24, dbydx=6, nbody=3, seed=31, re=30.0, precision=_abi.PIB_FP32_BULK),
             dict(kind="random", my=32, dbydx=8, nbody=4, seed=32, re=40.0, precision=_abi.PIB_FP64),
             dict(kind="random", my=20, dbydx=5, nbody=2, seed=33, re=20.0, precision=_abi.PIB_FP32_BULK),
             dict(kind="random", my=32, dbydx=8, nbody=2, seed=34, re=20.0, precision=_abi.PIB_FP64,
                  discr_scheme="upwind")]
    sts = [make_setup(tmp_path / f"c{q}", f"G{q}", **kw) for q, kw in enumerate(cases)]
    alone = [solver.Solver(st) for st in sts]
    batched = [solver.Solver(st) for st in sts]
    print(f"batch of {len(sts)} contexts, 7 steps", flush=True)
    B = solver.Batch(batched)
    B.step(3)
    B.step(4)
    for g in alone:
        g.step(7)
    for q, (a, b) in enumerate(zip(alone, batched)):
        _same(a, b, f"context {q} after 7 steps")
    # a moved bed in one context: its graphs are dropped and the batch graph rebuilt
    st = sts[0]
    xc = st.xc.copy()
    xc[0, 0] += 0.37 * st.config.dx
    for g in (alone[0], batched[0]):
        g.update_particles(xc)
    print("batch across a change of particles, 5 steps", flush=True)
    B.step(5)
    for g in alone:
        g.step(5)
    for q, (a, b) in enumerate(zip(alone, batched)):
        _same(a, b, f"context {q} after 12 steps")
    B.close()
    for g in alone + batched:
        g.close()


def test_batch_of_c3_realizations_against_the_oracle(tmp_path, orc):
    sts = [solver.RunSetup(configs.c3_realization(str(tmp_path), k)) for k in (0, 3)]
    gs = [solver.Solver(st) for st in sts]
    os_ = [solver.Solver(st, lib=orc.load(), prefix="orc_") for st in sts]
    print("C3 realizations 0 and 3 (96^3), FP64, batch of 2 vs oracle, 4 steps", flush=True)
    B = solver.Batch(gs)
    for n in range(4):
        B.step(1)
        for g, o, st in zip(gs, os_, sts):
            o.step(1)
            eu = rel_err(g.field("u"), o.field("u"))
            ep = rel_err(g.field("p"), o.field("p"))
            ef = rel_err(g.forces()["force"], o.forces()["force"])
            print(f"  step {n + 1} {st.nx}^3/{st.nbody}: u {eu:.2e} p {ep:.2e} force {ef:.2e}", flush=True)
            assert eu < 1e-10 and ep < 1e-10 and ef < 1e-9
    B.close()
    for g in gs + os_:
        g.close()


def test_autotune_keeps_the_state_and_caches_the_pick(tmp_path, orc):
    st = make_setup(tmp_path / "a", "T0", kind="random", my=32, dbydx=8, nbody=3, seed=41, re=30.0)
    g = solver.Solver(st)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    g.step(3)
    o.step(3)
    before = {f: g.field(f) for f in ("u", "p", "nl")}
    sc0 = g.scalars()
    picks, ms = g.autotune(nsteps=4, force=True)
    print(f"autotune 32^3: picks {picks}, {ms:.4f} ms/step", flush=True)
    assert picks[0] in (0, 1) and picks[1] in (0, 1, 2) and picks[2] in (0, 1) and ms > 0
    for f, a in before.items():
        assert np.array_equal(g.field(f), a), f
    sc1 = g.scalars()
    assert sc1.t == sc0.t and sc1.dt == sc0.dt and sc1.norm_drag == sc0.norm_drag
    nl = g.tuning(_abi.PIB_TUNE_NL)
    assert nl in (0, 1)
    assert (g.tuning(_abi.PIB_TUNE_FFT), g.tuning(_abi.PIB_TUNE_VELSTEP), g.tuning(_abi.PIB_TUNE_STREAMS)) == picks
    g2 = solver.Solver(st)
    assert g2.lib.pib_set_tuning(g2.ctx, _abi.PIB_TUNE_NL, 1 - nl) == 0 and g2.tuning(_abi.PIB_TUNE_NL) == 1 - nl
    picks2, ms2 = g2.autotune(nsteps=4)
    assert picks2 == picks and ms2 == 0.0 and g2.tuning(_abi.PIB_TUNE_NL) == nl
    g2.close()
    g.step(3)
    o.step(3)
    eu = rel_err(g.field("u"), o.field("u"))
    ef = rel_err(g.forces()["force"], o.forces()["force"])
    print(f"  tuned context after 3 more steps vs oracle: u {eu:.2e} force {ef:.2e}", flush=True)
    assert eu < 1e-10 and ef < 1e-9
    g.close()
    o.close()
