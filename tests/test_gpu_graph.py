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
