"""Independent contexts launched together (pib_step_async on each, then
pib_sync) give bitwise the same state as the same contexts stepped one after
the other: the contexts share nothing on the device."""
import numpy as np
import pytest

from conftest import make_setup
from pureibm_ps2d_fd1d_amd import _abi, solver

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
