"""GPU parity of the soft-sphere DEM (include/pureibm_dem.h) against the oracle
restatement of des_time_march (oracle/pib_dem_oracle.c), on the same seeded
beds: positions, velocities, IBMUPDATE outputs, contact forces and the
contact diagnostics after each march.  The GPU path sums every particle's
contact force in the reference's order with FP contraction off and the
device's correctly rounded sqrt and division, so positions and velocities
must agree bit for bit; the outputs derived from them are held to 1e-13
relative to the box."""
import numpy as np
import pytest

from pureibm_ps2d_fd1d_amd import _abi, solver
from conftest import make_setup, rel_err
from dem_beds import random_bed

pytestmark = pytest.mark.gpu

TOL = 1e-13


def pair(orc, xc, rad, vel, **kw):
    g = solver.Dem(xc, rad, velbdy=vel, **kw)
    o = solver.Dem(xc, rad, velbdy=vel, lib=orc.load(), prefix="orc_", **kw)
    return g, o


def compare(g, o, out_g, out_o, box):
    for a, b in zip(out_g, out_o):
        assert np.max(np.abs(a - b)) <= TOL * box
    sg, so = g.state(), o.state()
    for a, b in zip(sg[:4], so[:4]):
        assert np.max(np.abs(a - b)) <= TOL * max(1.0, np.max(np.abs(b)))
    ig, io = sg[4], so[4]
    assert ig.ncontacts == io.ncontacts and ig.factor == io.factor and ig.substeps == io.substeps
    assert abs(ig.overlap_max - io.overlap_max) <= 1e-9 * max(io.overlap_max, 1.0)
    assert ig.dtsolid == io.dtsolid and ig.tcoll == io.tcoll
    assert list(ig.cells) == list(io.cells) and list(ig.box) == list(io.box)
    return all(np.array_equal(a, b) for a, b in zip(sg[:2], so[:2]))


CASES = [
    # n, box, dx, radii (grid units), only_dem, seed
    (300, 8.0, 1.0 / 16, (8.0, 8.0), True, 11),
    (250, 8.0, 1.0 / 16, (6.0, 10.0), False, 12),
    (40, 4.6, 1.0 / 16, (6.0, 9.0), True, 13),        # two cells a direction: repeated images
    (1500, 14.0, 1.0 / 16, (7.0, 9.0), False, 14),    # three particles a thread
    (1200, 12.8, 1.0 / 20, (10.0, 10.0), False, 15),  # 256^3 bed of d = 20 cells
]


@pytest.mark.parametrize("case", CASES, ids=lambda c: f"n{c[0]}_box{c[1]}")
def test_dem_matches_oracle(orc, case):
    n, box, dx, rr, only_dem, seed = case
    print(f"DEM parity n={n} box={box} seed={seed}", flush=True)
    xc, rad, vel = random_bed(n, box, dx=dx, rad_range=rr, overlap=0.01, vel=1.0, seed=seed)
    kw = dict(dx=dx, xlength=box, ylength=box, zlength=box, coeff_rest=0.8, only_dem=only_dem,
              rhof=1.2, rhos=2.5, foffset=0.0)
    g, o = pair(orc, xc, rad, vel, **kw)
    rng = np.random.default_rng(seed)
    dts = g.state()[4].dtsolid
    exact = True
    for k, mult in enumerate([20, 20, 0.5, 37, 20, 60]):
        hyd = {}
        if not only_dem:
            hyd = dict(pres=rng.normal(0, 0.05, (n, 3)), visc=rng.normal(0, 0.05, (n, 3)),
                       mpg=rng.normal(0, 0.1, 3), frame_accln=rng.normal(0, 0.1, 3))
        out_g = g.march(mult * dts, **hyd)
        out_o = o.march(mult * dts, **hyd)
        exact &= compare(g, o, out_g, out_o, box)
    print(f"bitwise equal: {exact}", flush=True)
    assert exact
    g.close()
    o.close()


def test_gpu_head_on_restitution(orc):
    dx = 1.0 / 16
    xc = np.array([[1.5 / dx + 1, 2.5 / dx + 1, 2.5 / dx + 1], [2.51 / dx + 1, 2.5 / dx + 1, 2.5 / dx + 1]])
    vel = np.array([[1.0, 0.0, 0.0], [-1.0, 0.0, 0.0]])
    e = 0.6
    g = solver.Dem(xc, np.array([8.0, 8.0]), dx=dx, xlength=5.0, ylength=5.0, zlength=5.0, velbdy=vel,
                   coeff_rest=e, only_dem=True)
    info = g.state()[4]
    # the whole collision inside one march: separation 0.01 at closing speed 2
    g.march(0.005 + 2 * info.tcoll)
    v = g.state()[1]
    assert abs((v[1, 0] - v[0, 0]) / 2.0 - e) < 0.01
    assert v[0, 0] + v[1, 0] == 0.0
    g.close()


def test_gpu_motion_error_is_reported():
    dx = 1.0 / 16
    xc, rad, vel = random_bed(10, 6.0, seed=7)
    fast = vel * 0
    fast[0, 0] = 1e5
    g = solver.Dem(xc, rad, dx=dx, xlength=6.0, ylength=6.0, zlength=6.0, velbdy=fast, only_dem=True)
    with pytest.raises(solver.PibError) as ei:
        g.march(g.state()[4].dtsolid)
    assert ei.value.code == _abi.PIB_ERR_MOTION
    g.close()


def test_gpu_set_state_restarts_the_sub_cycle(orc):
    """A march after set_state equals a fresh context created at that state."""
    dx = 1.0 / 16
    xc, rad, vel = random_bed(200, 8.0, seed=21, vel=1.0)
    kw = dict(dx=dx, xlength=8.0, ylength=8.0, zlength=8.0, only_dem=True)
    g = solver.Dem(xc, rad, velbdy=vel, **kw)
    dts = g.state()[4].dtsolid
    g.march(30 * dts)
    pos, v = g.state()[:2]
    h = solver.Dem(xc, rad, velbdy=vel, **kw)
    o = solver.Dem(xc, rad, velbdy=vel, lib=orc.load(), prefix="orc_", **kw)
    for d in (g, h, o):
        d.set_state(pos, v)
    a, b, c = g.march(25 * dts), h.march(25 * dts), o.march(25 * dts)
    for x, y, z in zip(a, b, c):
        assert np.array_equal(x, y)
        assert np.max(np.abs(x - z)) <= TOL * 8.0
    for d in (g, h, o):
        d.close()


def test_dem_moves_the_bed_of_a_fluid_step(orc, tmp_path):
    """The chain bcset runs for moving particles: DEM march with the drag of
    the step, IBMUPDATE, the new solid mask and reversal pairs, then further
    fluid steps with moving surfaces -- GPU against the oracle (FP64)."""
    st = make_setup(tmp_path, "MV", kind="random", my=32, dbydx=8, nbody=6, seed=3)
    g = solver.Solver(st)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    for s_ in (g, o):
        s_.step(2)
    dx = st.config.dx
    L = (st.config.nx * dx, st.config.my * dx, st.config.mz * dx)
    rng = np.random.default_rng(5)
    vel0 = rng.uniform(-0.3, 0.3, (st.nbody, 3))
    kw = dict(dx=dx, xlength=L[0], ylength=L[1], zlength=L[2], coeff_rest=0.8, rhos=2.0, rhof=1.0,
              dia_phys=st.config.char_length)
    dg = solver.Dem(st.xc, st.radbdy, velbdy=vel0, **kw)
    do = solver.Dem(st.xc, st.radbdy, velbdy=vel0, lib=orc.load(), prefix="orc_", **kw)
    for it in range(2):
        fg, fo = g.forces(), o.forces()
        sg, so = g.scalars(), o.scalars()
        dt = sg.dt               # des_time_march(.FALSE.) advances the particles over the fluid dt
        outs = []
        for d, f, s_ in ((dg, fg, sg), (do, fo, so)):
            outs.append(d.march(dt, pres=f["pres"], visc=f["visc"], mpg=np.array(s_.mpg),
                                frame_accln=np.array(s_.frame_accln)))
        (xg, vg, ag), (xo, vo, ao) = outs
        assert rel_err(xg, xo) < 1e-12 and rel_err(vg, vo) < 1e-10
        cg, ng = g.update_particles(xg, vg, ag)
        co, no = o.update_particles(xg, vg, ag)
        assert cg == co and np.array_equal(ng, no)
        assert np.array_equal(g.field("mask"), o.field("mask"))
        for s_ in (g, o):
            s_.step(1)
        assert rel_err(g.field("u"), o.field("u")) < 1e-10
        assert abs(g.scalars().norm_drag - o.scalars().norm_drag) <= 1e-9 * abs(o.scalars().norm_drag)
    for x in (g, o, dg, do):
        x.close()


def test_moving_bed_steps_match_the_oracle(orc, tmp_path):
    """Freely moving bed, FP64: each step runs nonlinear to the drag, the DEM
    over dt with the moving-particle mpg and frame_accln, the new mask and
    reversal pairs, the forcing and velstep -- the GPU chain (GPU DEM) against
    the oracle chain (oracle DEM)."""
    st = make_setup(tmp_path, "MB", kind="random", my=32, dbydx=8, nbody=6, seed=3)
    st.config.move_particles = 1
    st.config.rhos, st.config.rhof = 2.0, 1.0
    g = solver.Solver(st)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    dx = st.config.dx
    L = (st.config.nx * dx, st.config.my * dx, st.config.mz * dx)
    vel0 = np.random.default_rng(5).uniform(-0.3, 0.3, (st.nbody, 3))
    kw = dict(dx=dx, xlength=L[0], ylength=L[1], zlength=L[2], coeff_rest=0.8, rhos=2.0, rhof=1.0,
              dia_phys=st.config.char_length)
    dg = solver.Dem(st.xc, st.radbdy, velbdy=vel0, **kw)
    do = solver.Dem(st.xc, st.radbdy, velbdy=vel0, lib=orc.load(), prefix="orc_", **kw)
    for s_ in (g, o):
        s_.update_particles(st.xc, vel0)
    solids = set()
    for k in range(5):
        g.step_moving(dg, 1)
        o.step_moving(do, 1)
        solids.add(g.count_solid)
        assert g.count_solid == o.count_solid and np.array_equal(g.nrpr_active, o.nrpr_active)
        assert rel_err(g.field("u"), o.field("u")) < 1e-10
        sg, so = g.scalars(), o.scalars()
        for name in ("mpg", "frame_accln", "frame_vel", "umean"):
            assert rel_err(np.array(getattr(sg, name)), np.array(getattr(so, name))) < 1e-9, name
        assert abs(sg.norm_drag - so.norm_drag) <= 1e-9 * abs(so.norm_drag)
        pg, po = dg.state()[0], do.state()[0]
        assert np.max(np.abs(pg - po)) < 1e-10 * L[0]
    assert any(np.array(sg.frame_accln) != 0.0)
    assert len(solids) > 1          # the mask changed as the bed moved
    for x in (g, o, dg, do):
        x.close()


def _moving(tmp_path, name, prec=_abi.PIB_FP64, rk3=False, lib=None, prefix="pib_", vel_seed=5, **kw):
    st = make_setup(tmp_path, name, kind="random", precision=prec, my=32, dbydx=8, nbody=6, seed=3,
                    imove=1, rhos=2.0, coeff_rest=0.8, rk3=rk3, **kw)
    assert st.config.move_particles == 1 and st.config.rhos == 2.0
    s = solver.Solver(st, lib=lib, prefix=prefix)
    dx = st.config.dx
    L = (st.config.nx * dx, st.config.my * dx, st.config.mz * dx)
    vel0 = np.random.default_rng(vel_seed).uniform(-0.3, 0.3, (st.nbody, 3))
    d = solver.Dem(st.xc, st.radbdy, velbdy=vel0, dx=dx, xlength=L[0], ylength=L[1], zlength=L[2],
                   coeff_rest=0.8, rhos=2.0, rhof=1.0, dia_phys=st.config.char_length, lib=lib,
                   prefix=prefix)
    s.update_particles(st.xc, vel0)
    return st, s, d


def test_moving_bed_rk3_matches_the_oracle(orc, tmp_path):
    """RK3: the DEM runs in every stage, as bcset calls it every stage."""
    st, g, dg = _moving(tmp_path, "MR", rk3=True)
    _, o, do = _moving(tmp_path, "MR", rk3=True, lib=orc.load(), prefix="orc_")
    for _ in range(2):
        g.step_moving(dg, 1)
        o.step_moving(do, 1)
        assert g.count_solid == o.count_solid
        assert rel_err(g.field("u"), o.field("u")) < 1e-10
        assert abs(g.scalars().norm_drag - o.scalars().norm_drag) <= 1e-9 * abs(o.scalars().norm_drag)
    assert dg.state()[4].substeps == do.state()[4].substeps > 0
    for x in (g, o, dg, do):
        x.close()


def test_moving_bed_with_the_scalar_matches_the_oracle(orc, tmp_path):
    """IMOVE = 1 with ISCALON = 1 (the reference's MIS deck combination):
    each stage ends with the scalar step on the moved bed's mask and
    reversal pairs; phi, the surface fluxes and the heat ratio follow the
    oracle chain (1e-10 on phi, 1e-9 on the fluxes)."""
    st, g, dg = _moving(tmp_path, "MS", iscalon=1, discr_scheme="upwind")
    _, o, do = _moving(tmp_path, "MS", lib=orc.load(), prefix="orc_", iscalon=1, discr_scheme="upwind")
    assert g.scalar_on and o.scalar_on
    for k in range(4):
        g.step_moving(dg, 1)
        o.step_moving(do, 1)
        e = rel_err(g.field("phi"), o.field("phi"))
        fg, fo = g.scalar_flux(), o.scalar_flux()
        ef = rel_err(fg[0], fo[0])
        print(f"  step {k + 1}: phi {e:.1e} flux {ef:.1e} heat ratio {g.scalar_get()[0]:.12f}", flush=True)
        assert e < 1e-10 and ef < 1e-9
        assert abs(g.scalar_get()[0] - o.scalar_get()[0]) <= 1e-12
    for x in (g, o, dg, do):
        x.close()


def test_moving_bed_fp32_bulk_drag_within_tolerance(tmp_path):
    """FP32-bulk fluid with the FP64 DEM: mean drag within 1e-4 of FP64."""
    _, g64, d64 = _moving(tmp_path, "M64")
    _, g32, d32 = _moving(tmp_path, "M32", prec=_abi.PIB_FP32_BULK)
    for _ in range(3):
        g64.step_moving(d64, 1)
        g32.step_moving(d32, 1)
    n64, n32 = g64.scalars().norm_drag, g32.scalars().norm_drag
    assert abs(n32 - n64) <= 1e-4 * abs(n64), (n32, n64)
    for x in (g64, g32, d64, d32):
        x.close()


def test_moving_bed_restart_continues(tmp_path):
    """A moving run saved and read back continues with the frame state and
    the centres of the save: the DEM starts again from XC/VELBDY, as the
    reference's des_time_march(.TRUE.) does after a restart, so the fields
    follow the uninterrupted run to round-off, not bit for bit."""
    st, a, da = _moving(tmp_path, "MS")
    a.step_moving(da, 2)
    pos, vel = da.state()[:2]
    xc = np.asfortranarray(pos / st.config.dx + 1.0)  # IBMUPDATE of the current state
    a.save_restart(xc=xc, velbdy=vel)
    sa = a.scalars()
    b = solver.Solver(st)
    cnt, _, xr, vr = b.load_restart()
    assert np.array_equal(xr, xc) and np.array_equal(vr, vel)
    sb = b.scalars()
    assert list(sb.frame_vel) == list(sa.frame_vel) and list(sb.frame_pos) == list(sa.frame_pos)
    dx = st.config.dx
    L = (st.config.nx * dx, st.config.my * dx, st.config.mz * dx)
    db = solver.Dem(xr, st.radbdy, velbdy=vr, dx=dx, xlength=L[0], ylength=L[1], zlength=L[2], coeff_rest=0.8,
                    rhos=2.0, rhof=1.0, dia_phys=st.config.char_length)
    a.step_moving(da, 2)
    b.step_moving(db, 2)
    assert rel_err(b.field("u"), a.field("u")) < 1e-9
    for x in (a, b, da, db):
        x.close()


def test_moving_particle_cfl_branch_matches_the_oracle(orc, tmp_path):
    """compute_new_timestep's moving-particle branch (dependent_functions.F90:
    1156-1165) on the GPU against the oracle: the fastest particle breaks the
    CFL bound while the fluid keeps it, so the new dt comes from the branch
    (cfl_max*dx/umax_tmp); one whole step then agrees with the oracle."""
    st = make_setup(tmp_path, "MC", kind="random", my=24, dbydx=6, nbody=4, seed=12, re=20.0, adaptive=True)
    st.config.move_particles = 1
    st.config.rhos, st.config.rhof = 2.0, 1.0
    g = solver.Solver(st)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    s0 = o.scalars()
    cfl_max, dx = st.config.cfl_max, st.config.dx
    up = 4.0 * cfl_max * dx / s0.dt
    vel = np.zeros((st.nbody, 3))
    vel[2] = [0.6 * up, 0.0, 0.8 * up]
    vel[0] = [0.1 * up, 0.1 * up, 0.0]
    for s_ in (g, o):
        s_.update_particles(st.xc, vel)
        s_.compute_new_timestep(1)
    sg, so = g.scalars(), o.scalars()
    assert so.umax_tmp * s0.dt / dx <= cfl_max and so.dt == cfl_max * dx / so.umax_tmp
    assert abs(sg.dt - so.dt) <= 1e-14 * so.dt and abs(sg.t - so.t) <= 1e-14 * so.t
    for s_ in (g, o):
        s_.nonlinear(1)
        s_.velstep(1)
    assert rel_err(g.field("u"), o.field("u")) < 1e-10
    assert rel_err(g.forces()["force"], o.forces()["force"]) < 1e-9
    g.close()
    o.close()
