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
