"""Whole steps replayed from HIP graphs against the same steps launched kernel
by kernel: bitwise equal after every replay batch, with eager work and host
synchronisation between the replays, for both nl/onl buffer assignments and
for the AB2 and RK3 step shapes."""
import numpy as np
import pytest

from conftest import make_setup
from pureibm_ps2d_fd1d_amd import _abi, solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [_abi.PIB_FP64, _abi.PIB_FP32_BULK])
@pytest.mark.parametrize("rk3", [False, True])
def test_graph_replays_equal_eager_steps(tmp_path, prec, rk3):
    print(f"graph replay prec={prec} rk3={rk3}", flush=True)
    kw = dict(kind="random", my=24, dbydx=6, nbody=3, seed=7, re=30.0, rk3=rk3)
    st = make_setup(tmp_path, "GR", precision=prec, **kw)
    ge = solver.Solver(st)
    gg = solver.Solver(st)
    assert ge.lib.pib_set_tuning(ge.ctx, _abi.PIB_TUNE_GRAPH, 0) == 0
    assert gg.lib.pib_set_tuning(gg.ctx, _abi.PIB_TUNE_GRAPH, 1) == 0
    for batch in (3, 4, 1, 5, 2):
        ge.step(batch)
        gg.step(batch)
        # host reads between replays (synchronises the stream)
        for f in ("u", "p", "nl", "ubcp"):
            a, b = ge.field(f), gg.field(f)
            assert np.array_equal(a, b), (batch, f)
        assert np.array_equal(ge.forces()["force"], gg.forces()["force"])
        se, sg = ge.scalars(), gg.scalars()
        for k in ("t", "dt", "norm_drag", "divmax"):
            assert getattr(se, k) == getattr(sg, k), (batch, k)
        assert list(se.mpg) == list(sg.mpg)
    ge.close()
    gg.close()


@pytest.mark.parametrize("prec", [_abi.PIB_FP64, _abi.PIB_FP32_BULK])
def test_forked_branches_equal_one_stream(tmp_path, prec):
    """The form_nl branch on its own stream gives the same bits as one stream,
    launched eagerly and replayed from graphs."""
    print(f"forked branches prec={prec}", flush=True)
    kw = dict(kind="random", my=24, dbydx=6, nbody=3, seed=11, re=30.0)
    st = make_setup(tmp_path, "FK", precision=prec, **kw)
    g1 = solver.Solver(st)
    g2 = solver.Solver(st)
    assert g1.lib.pib_set_tuning(g1.ctx, _abi.PIB_TUNE_STREAMS, 0) == 0
    assert g2.lib.pib_set_tuning(g2.ctx, _abi.PIB_TUNE_STREAMS, 1) == 0
    for batch in (1, 4, 3):
        g1.step(batch)
        g2.step(batch)
        for f in ("u", "p", "nl", "nlbcp", "fr", "ppr"):
            assert np.array_equal(g1.field(f), g2.field(f)), (batch, f)
        assert g1.scalars().norm_drag == g2.scalars().norm_drag
    g1.close()
    g2.close()


@pytest.mark.parametrize("prec", [_abi.PIB_FP64, _abi.PIB_FP32_BULK])
@pytest.mark.parametrize("scheme,my,dbydx", [("center", 24, 6), ("upwind", 32, 8), ("center", 128, 16)])
def test_nl_by_separate_transforms_equals_in_place(tmp_path, prec, scheme, my, dbydx):
    """form_nl as eight separate transforms a plane (PIB_TUNE_NL = 1, nl
    formed once in the qin transform) against the in-place three-stage form
    (PIB_TUNE_NL = 0): the same transforms of the same inputs summed in the
    same order, on the generic and the compiled 128^2 plane.  Whole steps,
    graph-replayed; u, p, nl and the forces after each batch agree to 1e-11
    (FP64) / 1e-5 (FP32-bulk) of their largest entry (the two kernels may
    contract the flux arithmetic into fused multiply-adds differently, and
    the upwind x-flux difference of a nearly uniform stream cancels); the run
    prints the largest difference and whether they are bitwise equal."""
    print(f"nl forms prec={prec} scheme={scheme} my={my}", flush=True)
    kw = dict(kind="random", my=my, dbydx=dbydx, nbody=3, seed=13, re=30.0, discr_scheme=scheme)
    st = make_setup(tmp_path, "NLF", precision=prec, **kw)
    g0 = solver.Solver(st)
    g1 = solver.Solver(st)
    assert g0.lib.pib_set_tuning(g0.ctx, _abi.PIB_TUNE_NL, 0) == 0
    assert g1.lib.pib_set_tuning(g1.ctx, _abi.PIB_TUNE_NL, 1) == 0
    tol = 1e-11 if prec == _abi.PIB_FP64 else 1e-5
    for batch in (1, 3, 2):
        g0.step(batch)
        g1.step(batch)
        same, worst = True, 0.0
        for f in ("u", "p", "nl"):
            a, b = g0.field(f), g1.field(f)
            same = same and np.array_equal(a, b)
            # (nl of the first steps from the uniform stream is rounding
            # noise of a zero field: the scale floor keeps it from counting)
            e = np.max(np.abs(a - b)) / max(np.max(np.abs(a)), 1e-6)
            worst = max(worst, e)
            assert e <= tol, (batch, f, e)
        fa, fb = g0.forces()["force"], g1.forces()["force"]
        assert np.max(np.abs(fa - fb)) <= tol * 10 * np.max(np.abs(fa)), batch
        print(f"  batch {batch}: largest difference {worst:.1e}, bitwise equal {same and np.array_equal(fa, fb)}",
              flush=True)
    g0.close()
    g1.close()
