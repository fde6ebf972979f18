"""The same steps issued as one pib_step call, as single-step calls, and as
single-step calls with host reads between them: all equal to the oracle."""
import numpy as np
import pytest

from conftest import make_setup, rel_err
from pureibm_ps2d_fd1d_amd import solver

pytestmark = pytest.mark.gpu


def test_single_step_calls_match_one_call(tmp_path, orc):
    print("single-step calls vs one call, 24^3 bed", flush=True)
    kw = dict(kind="random", my=24, dbydx=6, nbody=3, seed=5, re=40.0)
    st = make_setup(tmp_path, "SC", **kw)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    g1 = solver.Solver(st)
    g2 = solver.Solver(st)
    g3 = solver.Solver(st)
    for n in range(1, 4):
        o.step(1)
        g2.step(1)
        g3.step(1)
        s3 = g3.scalars()
        f3 = g3.field("fr")
        uo = o.field("u")
        e2 = rel_err(g2.field("u"), uo)
        e3 = rel_err(g3.field("u"), uo)
        print(f"  step {n}: u err single {e2:.2e} single+reads {e3:.2e} max|fr| gpu {np.abs(f3).max():.3e} "
              f"oracle {np.abs(o.field('fr')).max():.3e} norm_drag {s3.norm_drag:.6e} {o.scalars().norm_drag:.6e}",
              flush=True)
    g1.step(3)
    e1 = rel_err(g1.field("u"), o.field("u"))
    print(f"  one call of 3: u err {e1:.2e}", flush=True)
    assert e1 < 1e-10
    assert rel_err(g2.field("u"), o.field("u")) < 1e-10
    assert rel_err(g3.field("u"), o.field("u")) < 1e-10
