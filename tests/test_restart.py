"""Restart files in the reference's wire format (restart_funcs.F90).

The record layout is pinned against the reference's WRITE statements
(save_unformatted :188-362): record by record, the byte lengths of a
sequential unformatted file with the usual 4-byte record markers, and the
values they carry.  The round trip is pinned by continuation: a run saved
after n steps, read into a fresh context and stepped m more equals the
uninterrupted n+m-step run.  No GPU here (oracle contexts)."""
import os
import struct

import numpy as np
import pytest

from conftest import make_setup, rel_err
from pureibm_ps2d_fd1d_amd import solver


def records(path):
    """Fortran sequential unformatted records: [len][payload][len] each."""
    out = []
    with open(path, "rb") as f:
        data = f.read()
    pos = 0
    while pos < len(data):
        (n,) = struct.unpack_from("<i", data, pos)
        payload = data[pos + 4:pos + 4 + n]
        (m,) = struct.unpack_from("<i", data, pos + 4 + n)
        assert m == n, path
        out.append(payload)
        pos += 8 + n
    return out


@pytest.fixture
def oracle_run(orc, tmp_path):
    st = make_setup(tmp_path, "RS", kind="random", my=32, dbydx=8, nbody=5, seed=2)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    return st, o


def test_records_follow_save_unformatted(oracle_run, tmp_path):
    st, o = oracle_run
    o.step(3)
    c = o.save_restart()
    assert c == 1
    run = st.run_name
    assert open(run + "_RESTART").read().split() == ["1"]
    nx, my2, mz, nb = o.nx, o.my2, o.mz, st.nbody
    s = o.scalars()
    from pureibm_ps2d_fd1d_amd import _abi
    nerr = _abi.load_host().pib_host_get_nsteps_hint()
    u = records(run + "_u_1.rst")
    assert [len(r) for r in u] == [16, 24, 24, 24, 8, 4, 16, 8 * nerr, 16, 8 * nerr, 16 * (nx + 1) * my2 * mz * 3]
    assert struct.unpack("<2d", u[0]) == (s.t, s.dt)
    assert struct.unpack("<3d", u[1]) == tuple(s.umean)
    assert struct.unpack("<d", u[4]) == (s.cf,)
    assert struct.unpack("<i", u[5]) == (nerr,)
    assert struct.unpack("<2d", u[6]) == (s.ferror, s.fold)
    assert u[-1] == np.asfortranarray(o.field("u")).tobytes(order="F")
    p = records(run + "_p_1.rst")
    assert [len(r) for r in p] == [24, 16 * nx * my2 * mz]
    assert struct.unpack("<3d", p[0]) == tuple(s.mpg)
    nl = records(run + "_nl_1.rst")
    assert nl == [np.asfortranarray(o.field("nl")).tobytes(order="F")]
    cfg = records(run + "_sphere_config.rst")
    assert [len(r) for r in cfg] == [4, 4, 4, 8, 24, 8, 24 * nb, 8 * nb, 24 * nb]
    assert struct.unpack("<i", cfg[0]) == (nb,)
    assert np.array_equal(np.frombuffer(cfg[6]).reshape((nb, 3), order="F"), st.xc)
    assert np.array_equal(np.frombuffer(cfg[7]), st.radbdy)


def test_count_alternates_and_old_files_go(oracle_run):
    st, o = oracle_run
    run = st.run_name
    assert o.save_restart() == 1
    o.step(1)
    assert o.save_restart() == 2
    assert os.path.exists(run + "_u_2.rst") and not os.path.exists(run + "_u_1.rst")
    assert open(run + "_RESTART").read().split() == ["2"]
    o.step(1)
    assert o.save_restart() == 1
    assert os.path.exists(run + "_u_1.rst") and not os.path.exists(run + "_u_2.rst")


def test_continuation_equals_the_uninterrupted_run(orc, tmp_path):
    st = make_setup(tmp_path, "RC", kind="random", my=32, dbydx=8, nbody=5, seed=4)
    a = solver.Solver(st, lib=orc.load(), prefix="orc_")
    a.step(3)
    saved = a.ferror_history()[0]
    a.save_restart()
    a.step(2)
    b = solver.Solver(st, lib=orc.load(), prefix="orc_")
    cnt, ferr, xc, vel = b.load_restart()
    # the drag-error history travels with the restart (restart_funcs.F90:
    # 270-274): three entries of the run (the first exactly 1, fold starting
    # at 0), the rest the initial ones
    assert cnt == 1 and np.array_equal(ferr, saved)
    assert ferr[2] == 1.0 and np.all(ferr[3:] == 1.0) and ferr[0] != 1.0
    b.step(2)
    for k in ("u", "p", "nl"):
        assert np.array_equal(b.field(k), a.field(k)), k
    sa, sb = a.scalars(), b.scalars()
    assert sa.t == sb.t and sa.norm_drag == sb.norm_drag
    ha, hb = a.ferror_history(), b.ferror_history()
    assert np.array_equal(ha[0], hb[0]) and ha[1] == hb[1]


def test_missing_restart_is_an_error(oracle_run):
    st, o = oracle_run
    with pytest.raises(solver.PibError):
        o.load_restart()


def _scal_steps(s, n):
    for _ in range(n):
        s.compute_new_timestep(1)
        s.nonlinear(1)
        s.velstep(1)
        s.scalstep(1)


def _with_iscal_restart(st):
    p = st.run_name + "_floparam.in"
    txt = open(p).read()
    assert "ISCAL_RESTART = 0" in txt
    open(p, "w").write(txt.replace("ISCAL_RESTART = 0", "ISCAL_RESTART = 1"))
    return solver.RunSetup(st.run_name)


def test_scalar_restart_records_and_continuation(orc, tmp_path):
    """<RUN>_scal_<c>.rst (restart_funcs.F90:315-333): nu_error, nu_old;
    the NERR_STEPS history; phif; nlphif(1:nx+1); phisurfall; phirmean,
    heat_ratio, heat_ratio_old.  Read back with ISCAL_RESTART = 1 into a
    fresh context, the scalar run continues bit for bit; the next save takes
    count 2 and removes the count-1 scalar file."""
    st = make_setup(tmp_path, "SR", kind="random", my=24, dbydx=6, nbody=3, seed=8, discr_scheme="upwind",
                    iscalon=1)
    a = solver.Solver(st, lib=orc.load(), prefix="orc_")
    _scal_steps(a, 3)
    c = a.save_restart()
    from pureibm_ps2d_fd1d_amd import _abi
    nerr = _abi.load_host().pib_host_get_nsteps_hint()
    nx, my2, mz, nb = a.nx, a.my2, a.mz, st.nbody
    r = records(st.run_name + f"_scal_{c}.rst")
    assert [len(x) for x in r] == [16, 8 * nerr, 16 * (nx + 1) * my2 * mz, 16 * (nx + 1) * my2 * mz, 8 * nb, 24]
    assert r[2] == np.asfortranarray(a.field("phi")).tobytes(order="F")
    assert r[3] == np.asfortranarray(a.field("nlphi")).tobytes(order="F")
    hr = a.scalar_get()
    assert struct.unpack("<3d", r[5]) == (0.0, hr[0], hr[2]) and hr[0] != 1.0
    assert struct.unpack(f"<{nb}d", r[4]) == (0.0,) * nb
    _scal_steps(a, 2)
    b = solver.Solver(_with_iscal_restart(st), lib=orc.load(), prefix="orc_")
    b.load_restart()
    _scal_steps(b, 2)
    for k in ("u", "phi", "nlphi", "onlphi", "ffphi"):
        assert np.array_equal(b.field(k), a.field(k)), k
    assert b.scalar_get()[:3] == a.scalar_get()[:3]
    assert b.save_restart() == 2
    assert os.path.exists(st.run_name + "_scal_2.rst") and not os.path.exists(st.run_name + "_scal_1.rst")
