"""Restart files on the GPU path: a GPU run saved in the reference's restart
format and read into a fresh GPU context continues bit for bit like the
uninterrupted run (FP64 and FP32-bulk); files written from a GPU context,
read into the oracle, continue within the FP64 parity tolerance (1e-10)."""
import numpy as np
import pytest

from conftest import make_setup, rel_err
from pureibm_ps2d_fd1d_amd import _abi, solver

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("prec", [_abi.PIB_FP64, _abi.PIB_FP32_BULK])
def test_gpu_continuation_equals_the_uninterrupted_run(tmp_path, prec):
    st = make_setup(tmp_path, "GR", kind="random", precision=prec, my=32, dbydx=8, nbody=5, seed=4)
    a = solver.Solver(st)
    a.step(4)
    assert a.save_restart() == 1
    a.step(3)
    b = solver.Solver(st)
    b.load_restart()
    b.step(3)
    for k in ("u", "p", "nl"):
        assert np.array_equal(b.field(k), a.field(k)), k
    assert a.scalars().norm_drag == b.scalars().norm_drag
    a.close()
    b.close()


def test_gpu_restart_files_continue_on_the_oracle(orc, tmp_path):
    st = make_setup(tmp_path, "GO", kind="random", my=32, dbydx=8, nbody=5, seed=6)
    g = solver.Solver(st)
    g.step(3)
    g.save_restart()
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    o.load_restart()
    g.step(2)
    o.step(2)
    assert rel_err(g.field("u"), o.field("u")) < 1e-10
    assert abs(g.scalars().norm_drag - o.scalars().norm_drag) <= 1e-9 * abs(o.scalars().norm_drag)
    g.close()
    o.close()


def test_gpu_scalar_restart_continues(tmp_path):
    """An ISCALON run saved with its scalar file and read back into a fresh
    GPU context continues like the uninterrupted run (to 1e-12: the
    scalar's spreading adds with atomics, so two GPU runs agree to
    round-off rather than bitwise)."""
    from test_restart import _scal_steps, _with_iscal_restart
    st = make_setup(tmp_path, "GS", kind="random", my=32, dbydx=8, nbody=4, seed=9, discr_scheme="upwind",
                    iscalon=1)
    a = solver.Solver(st)
    _scal_steps(a, 3)
    assert a.save_restart() == 1
    _scal_steps(a, 3)
    b = solver.Solver(_with_iscal_restart(st))
    b.load_restart()
    _scal_steps(b, 3)
    for k in ("phi", "nlphi"):
        e = rel_err(b.field(k), a.field(k))
        print(f"  {k} {e:.1e}", flush=True)
        assert e < 1e-12, k
    ra, rb = a.scalar_get(), b.scalar_get()
    assert all(abs(x - y) <= 1e-13 * abs(x) for x, y in zip(ra[:3], rb[:3]))
    a.close()
    b.close()
