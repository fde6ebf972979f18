"""GPU parity: the HIP path through the C-ABI against the CPU oracle.

Tolerances (FP64): relative 1e-10 on velocity/pressure fields and 1e-9 on
per-particle drag after N steps; transforms and tridiagonal solves to 1e-12.
"""
import ctypes

import numpy as np
import pytest

from conftest import make_setup, rel_err
from pureibm_ps2d_fd1d_amd import _abi, configs, solver

pytestmark = pytest.mark.gpu


def _cfg(my, mz, nx=4, prec=0):
    c = _abi.PibConfig()
    c.nx, c.my, c.mz, c.precision = nx, my, mz, prec
    c.dx, c.vis, c.cfl_max = 0.1, 0.01, 0.5
    c.dorpr, c.revernorm, c.soft_start = 1, 1, 1
    c.ucharmod, c.char_length, c.dia, c.dr, c.voldom = 1.0, 1.0, 1.0, 1.0, 1.0
    return c


def _ctx(cfg):
    lib = _abi.load_gpu()
    ctx = ctypes.c_void_p()
    st = lib.pib_create(ctypes.byref(cfg), ctypes.byref(ctx))
    return lib, ctx, st


@pytest.mark.parametrize("my,mz,prec", [(8, 8, 0), (16, 12, 0), (24, 20, 0), (64, 64, 0), (96, 96, 0),
                                         (128, 128, 0), (128, 128, 1), (160, 160, 1), (160, 160, 0),
                                         (256, 256, 1)])
def test_transform_layer(orc, my, mz, prec):
    """pib_ff2rc_planes / pib_ff2cr_planes against the oracle's ff2rc/ff2cr,
    with non-Hermitian DC/Nyquist input for c2r."""
    print(f"transform my={my} mz={mz} prec={prec}", flush=True)
    lib, ctx, st = _ctx(_cfg(my, mz, prec=prec))
    assert st == 0
    rng = np.random.default_rng(my * 1000 + mz)
    np_ = 3
    a = rng.standard_normal((np_, mz, my))
    b = np.zeros((np_, mz, my // 2 + 1), dtype=np.complex128)
    d = _abi.dptr
    assert lib.pib_ff2rc_planes(ctx, d(a), my * mz, np_, d(b.view(np.float64)), (my // 2 + 1) * mz) == 0
    tol = 1e-13 if prec == 0 else 2e-6
    for q in range(np_):
        ref = orc.ff2rc(a[q].T)  # (my2, mz)
        assert rel_err(b[q].T, ref) < tol
    c = rng.standard_normal((np_, mz, my // 2 + 1)) + 1j * rng.standard_normal((np_, mz, my // 2 + 1))
    out = np.zeros((np_, mz, my))
    assert lib.pib_ff2cr_planes(ctx, d(c.view(np.float64)), (my // 2 + 1) * mz, np_, d(out), my * mz) == 0
    for q in range(np_):
        ref = orc.ff2cr(c[q].T, my)
        assert rel_err(out[q].T, ref) < tol
    lib.pib_destroy(ctx)


def test_transform_size_not_235_smooth_is_an_error():
    """my/2 = 7 has no 2,3,5 factorisation: refused at create (planes too
    large for the LDS are not an error: they go to HBM slots)."""
    lib, ctx, st = _ctx(_cfg(14, 16, prec=0))
    assert st == _abi.PIB_ERR_UNSUPPORTED


@pytest.mark.parametrize("cyclic", [0, 1])
def test_tridiag(orc, cyclic):
    print(f"tridiag cyclic={cyclic}", flush=True)
    lib, ctx, st = _ctx(_cfg(8, 8))
    n, ns = 40, 300
    rng = np.random.default_rng(7 + cyclic)
    a = rng.uniform(0.5, 1.0, (n, ns))
    c = rng.uniform(0.5, 1.0, (n, ns))
    b = -(a + c) - rng.uniform(0.5, 2.0, (n, ns))
    r = rng.standard_normal((n, ns))
    x = np.zeros((n, ns))
    d = _abi.dptr
    assert lib.pib_tridiag_batched(ctx, n, ns, d(a), d(b), d(c), d(r), d(x), cyclic) == 0
    for s in range(0, ns, 37):
        if cyclic:
            ref, _ = orc.cyclic_solve(a[:, s], b[:, s], c[:, s], r[:, s])
        else:
            ref, _ = orc.tridag(a[:, s], b[:, s], c[:, s], r[:, s])
        assert rel_err(x[:, s], ref) < 1e-12
    lib.pib_destroy(ctx)


def _pair(tmp_path, orc, name, kind="simple", precision=0, **kw):
    st = make_setup(tmp_path, name, kind=kind, precision=precision, **kw)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    g = solver.Solver(st)
    return st, o, g


def test_setup_mask_and_reversal_pairs(tmp_path, orc):
    print("setup mask/if_rev", flush=True)
    st, o, g = _pair(tmp_path, orc, "M", kind="random", my=32, dbydx=8, nbody=12, seed=3, re=20.0)
    assert g.count_solid == o.count_solid
    assert np.array_equal(g.nrpr_active, o.nrpr_active)
    assert np.array_equal(g.field("mask"), o.field("mask"))


def _advance_oracle_state(o, g, nsteps):
    """Advance the oracle, then copy its state (u, p, nl, scalars) to the GPU."""
    o.step(nsteps)
    for f in ("u", "p", "nl"):
        g.set_field(f, o.field(f))
    g.set_scalars(o.scalars())


@pytest.mark.parametrize("scheme", ["center", "upwind"])
def test_routines_one_step(tmp_path, orc, scheme):
    """Routine-by-routine comparison of one step from an oracle-advanced state."""
    print(f"routines scheme={scheme}", flush=True)
    st, o, g = _pair(tmp_path, orc, "R" + scheme, kind="random", my=24, dbydx=6, nbody=3, seed=11,
                     re=30.0, discr_scheme=scheme)
    _advance_oracle_state(o, g, 4)
    o.compute_new_timestep(1)
    g.compute_new_timestep(1)
    so, sg = o.scalars(), g.scalars()
    for k in ("dt", "t", "cf", "umax_tmp"):
        assert abs(getattr(sg, k) - getattr(so, k)) <= 1e-12 * abs(getattr(so, k)), k
    assert rel_err(g.field("ubcp"), o.field("ubcp")) < 1e-12
    assert rel_err(g.field("pbcp"), o.field("pbcp")) < 1e-12
    assert rel_err(np.array(sg.ufmean), np.array(so.ufmean)) < 1e-12
    o.nonlinear(1)
    g.nonlinear(1)
    for f in ("nl", "nlbcp", "onlbcp", "diffn", "omega", "ppr", "fr", "ff"):
        e = rel_err(g.field(f), o.field(f))
        print(f"  {f}: {e:.3e}", flush=True)
        assert e < 1e-11, f
    fo, fg = o.forces(), g.forces()
    for k in ("force", "pres", "visc", "torq"):
        assert rel_err(fg[k], fo[k]) < 1e-11, k
    so, sg = o.scalars(), g.scalars()
    for k in ("norm_drag", "ferror"):
        assert abs(getattr(sg, k) - getattr(so, k)) <= 1e-10 * abs(getattr(so, k)), k
    for k in ("mpg", "frmean", "pres_total", "visc_total"):
        assert rel_err(np.array(getattr(sg, k)), np.array(getattr(so, k))) < 1e-10, k
    o.velstep(1)
    g.velstep(1)
    for f in ("u", "p"):
        e = rel_err(g.field(f), o.field(f))
        print(f"  {f}: {e:.3e}", flush=True)
        assert e < 1e-11, f
    so, sg = o.scalars(), g.scalars()
    assert rel_err(np.array(sg.umean), np.array(so.umean)) < 1e-12


CASES = {
    "simple_center": dict(kind="simple", my=16, dbydx=6, re=20.0),
    "bed_upwind": dict(kind="random", my=24, dbydx=6, nbody=3, seed=5, re=40.0, discr_scheme="upwind"),
    "bed_minmod_rk3": dict(kind="random", my=20, dbydx=6, nbody=2, seed=9, re=30.0, discr_scheme="upwind",
                           limiter="minmod", rk3=True),
    "ramp_revernorm0": dict(kind="random", my=24, dbydx=8, nbody=2, seed=2, re=20.0, ramp_frac_steps=5,
                            revernorm=0, include_frmeanfluid=True),
    "set_mpg": dict(kind="random", my=24, dbydx=8, nbody=2, seed=4, re=20.0, set_mpg=True,
                    mpg=(-0.02, 0.0, 0.0)),
    # calc_bnd_forcing beside the reversal pairs (the reference's compiled
    # default DOBND = T), once on the LDS plane path and once on a compiled shape
    "dobnd": dict(kind="random", my=24, dbydx=6, nbody=3, seed=7, re=30.0, dobnd=True, discr_scheme="upwind"),
    "dobnd_frmeanfluid_32": dict(kind="random", my=32, dbydx=8, nbody=2, seed=3, re=20.0, dobnd=True,
                                 include_frmeanfluid=True),
    "dobnd_only": dict(kind="random", my=24, dbydx=8, nbody=2, seed=6, re=20.0, dobnd=True, dorpr=False),
}


def _start_moving(*solvers, u0=0.2):
    """SET_MPG runs start with umean = 0; give them a mean flow so the drag
    normalisation by the mean slip is defined."""
    for sv in solvers:
        s = sv.scalars()
        s.umean[0] = u0
        sv.set_scalars(s)


@pytest.mark.parametrize("case", list(CASES))
def test_multistep(tmp_path, orc, case):
    kw = dict(CASES[case])
    nsteps = 6 if kw.get("rk3") else 12
    print(f"multistep {case} nsteps={nsteps}", flush=True)
    st, o, g = _pair(tmp_path, orc, case, **kw)
    if kw.get("set_mpg"):
        _start_moving(o, g)
    o.step(nsteps)
    g.step(nsteps)
    eu = rel_err(g.field("u"), o.field("u"))
    ep = rel_err(g.field("p"), o.field("p"))
    ef = rel_err(g.forces()["force"], o.forces()["force"])
    print(f"  u {eu:.2e} p {ep:.2e} force {ef:.2e}", flush=True)
    assert eu < 1e-10 and ep < 1e-10
    assert ef < 1e-9
    so, sg = o.scalars(), g.scalars()
    assert abs(sg.norm_drag - so.norm_drag) <= 1e-9 * abs(so.norm_drag)
    assert abs(sg.t - so.t) <= 1e-12 * abs(so.t)


def test_c1_size_parity(tmp_path, orc):
    """64^3 single sphere of 16 cells (the simple-cubic workload) for a few steps."""
    print("c1-size parity", flush=True)
    st, o, g = _pair(tmp_path, orc, "C1", kind="simple", my=64, dbydx=16, re=20.0)
    o.step(3)
    g.step(3)
    eu = rel_err(g.field("u"), o.field("u"))
    ef = rel_err(g.forces()["force"], o.forces()["force"])
    print(f"  u {eu:.2e} force {ef:.2e}", flush=True)
    assert eu < 1e-10 and ef < 1e-9


def test_fp32_bulk_drag_error(tmp_path, orc):
    """FP32-bulk against the FP64 oracle: mean drag within 1e-4."""
    kw = dict(kind="random", my=32, dbydx=8, nbody=6, seed=21, re=50.0)
    print("fp32-bulk drag", flush=True)
    st = make_setup(tmp_path, "F32", **kw)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    st32 = make_setup(tmp_path, "F32", precision=_abi.PIB_FP32_BULK, **kw)
    g = solver.Solver(st32)
    o.step(20)
    g.step(20)
    so, sg = o.scalars(), g.scalars()
    err = abs(sg.norm_drag - so.norm_drag) / abs(so.norm_drag)
    print(f"  norm_drag fp64 {so.norm_drag:.8f} fp32 {sg.norm_drag:.8f} rel {err:.2e}", flush=True)
    assert err < 1e-4


def test_blowup_detected_like_the_reference(tmp_path, orc):
    """SET_MPG from rest: the drag normalisation by a zero mean slip makes
    norm_drag exceed 1e6 in the first step (boundary_condition.F90:783-884);
    both paths report the blow-up instead of stopping."""
    st, o, g = _pair(tmp_path, orc, "BU", kind="random", my=24, dbydx=8, nbody=2, seed=4, re=20.0,
                     set_mpg=True, mpg=(-0.02, 0.0, 0.0))
    with pytest.raises(solver.PibError) as eo:
        o.step(1)
    with pytest.raises(solver.PibError) as eg:
        g.step(1)
    assert eo.value.code == _abi.PIB_ERR_BLOWUP == eg.value.code


@pytest.mark.parametrize("my,dbydx,nbody,prec", [(32, 8, 6, 0), (64, 16, 3, 0), (128, 16, 8, 0), (96, 12, 6, 0),
                                                 (96, 12, 6, 1), (160, 20, 5, 1)])
def test_compiled_plane_shapes_match_generic_and_oracle(tmp_path, orc, my, dbydx, nbody, prec):
    """The two-stage plane transforms of the compiled shapes (my/2, mz) =
    (16,32), (32,64), (64,128), (48,96), (80,160) -- the last two with the
    radix-6/12 and radix-10 splits -- against the generic mixed-radix
    transforms (PIB_TUNE_FFT = 0) and against the oracle, over whole steps.
    FP64: 1e-10 on u, 1e-9 on the forces, compiled vs generic nl 1e-12.
    FP32-bulk (the 160^2 FP64 plane does not fit the LDS): 1e-5, 1e-4, 1e-5."""
    print(f"compiled shapes my={my} prec={prec}", flush=True)
    kw = dict(kind="random", my=my, dbydx=dbydx, nbody=nbody, seed=31, re=30.0, precision=prec)
    st = make_setup(tmp_path, "SH", **kw)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    g1 = solver.Solver(st)
    g0 = solver.Solver(st)
    assert g0.lib.pib_set_tuning(g0.ctx, _abi.PIB_TUNE_FFT, 0) == 0
    nsteps = 2 if my == 128 else 4
    o.step(nsteps)
    g1.step(nsteps)
    g0.step(nsteps)
    uo = o.field("u")
    e1, e0 = rel_err(g1.field("u"), uo), rel_err(g0.field("u"), uo)
    f1 = rel_err(g1.forces()["force"], o.forces()["force"])
    en = rel_err(g1.field("nl"), g0.field("nl"))
    print(f"  u err compiled {e1:.2e} generic {e0:.2e} force {f1:.2e} nl compiled vs generic {en:.2e}", flush=True)
    tu, tf, tn = (1e-10, 1e-9, 1e-12) if prec == 0 else (1e-5, 1e-4, 1e-5)
    assert e1 < tu and e0 < tu and f1 < tf
    assert en < tn


@pytest.mark.parametrize("my,dbydx,nbody,scheme", [(24, 6, 3, "center"), (40, 10, 2, "upwind")])
def test_planes_in_hbm_match_lds_and_oracle(tmp_path, orc, my, dbydx, nbody, scheme):
    """Plane transforms whose buffer is a slot of HBM scratch (the path of planes
    too large for the LDS of a workgroup, forced here with PIB_TUNE_FFT = 2,
    including the in-place generic form_nl) against the LDS-resident generic
    transforms and against the oracle, over whole steps."""
    print(f"planes in HBM my={my} scheme={scheme}", flush=True)
    kw = dict(kind="random", my=my, dbydx=dbydx, nbody=nbody, seed=41, re=30.0, discr_scheme=scheme)
    st = make_setup(tmp_path, "HB", **kw)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    g2 = solver.Solver(st)
    g0 = solver.Solver(st)
    assert g2.lib.pib_set_tuning(g2.ctx, _abi.PIB_TUNE_FFT, 2) == 0
    assert g0.lib.pib_set_tuning(g0.ctx, _abi.PIB_TUNE_FFT, 0) == 0
    o.step(4)
    g2.step(4)
    g0.step(4)
    uo = o.field("u")
    e2, e0 = rel_err(g2.field("u"), uo), rel_err(g0.field("u"), uo)
    f2 = rel_err(g2.forces()["force"], o.forces()["force"])
    print(f"  u err hbm {e2:.2e} lds {e0:.2e} force {f2:.2e}", flush=True)
    assert e2 < 1e-10 and e0 < 1e-10 and f2 < 1e-9
    assert rel_err(g2.field("nl"), g0.field("nl")) < 1e-12
    assert rel_err(g2.field("ubcp"), g0.field("ubcp")) < 1e-12
    o.close()
    g2.close()
    g0.close()


def test_c5_size_bed_runs_and_precisions_agree(tmp_path):
    """The 256^3 box of the largest workload (1 200 spheres of 20 cells, fixed
    bed): planes too large for the LDS run in HBM slots.  No oracle run at this
    size; the size-independent checks are the discrete divergence after the
    projection (round-off) and FP32-bulk against FP64."""
    print("c5-size bed 256^3", flush=True)
    d = str(tmp_path)
    configs.random_bed(d, "C5F", my=256, dbydx=20, nbody=1200, seed=5, re=50.0)
    runs = {}
    for prec in (_abi.PIB_FP64, _abi.PIB_FP32_BULK):
        st = solver.RunSetup(d + "/C5F", precision=prec)
        g = solver.Solver(st)
        g.step(3)
        runs[prec] = g.scalars()
        g.close()
    s64, s32 = runs[_abi.PIB_FP64], runs[_abi.PIB_FP32_BULK]
    print(f"  norm_drag fp64 {s64.norm_drag:.8e} fp32 {s32.norm_drag:.8e} divmax {s64.divmax:.2e}", flush=True)
    assert s64.divmax < 1e-10
    assert abs(s32.norm_drag - s64.norm_drag) < 1e-4 * abs(s64.norm_drag)


@pytest.mark.parametrize("rk3", [False, True])
def test_ferror_history_through_graph_replayed_steps(tmp_path, orc, rk3):
    """bcset's drag-error history (boundary_condition.F90:828-833): the GPU
    keeps it as a ring that the drag scalars write once a step, the oracle
    shifts the reference's array.  Seven steps in two pib_step calls (steps 3
    to 6 replay the captured graph) over a 3-entry history that starts from
    given values, so the history wraps twice; with RK3 only the last stage's
    ferror enters."""
    print(f"ferror history, 24^3 random bed, rk3={rk3}", flush=True)
    st = make_setup(tmp_path, "FH", kind="random", my=24, dbydx=6, nbody=3, seed=4, re=30.0, rk3=rk3)
    g = solver.Solver(st)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    a0, h0 = g.ferror_history()
    assert (a0 == 1.0).all() and h0 == 1.0  # NERR_STEPS entries of 1 from creation
    start = [0.5, 0.25, 0.125]
    for b in (g, o):
        b.set_ferror_history(3, start)
    ag, hg = g.ferror_history(3)
    assert list(ag) == start and hg == o.ferror_history(3)[1]
    g.step(2)
    g.step(5)
    o.step(7)
    ag, hg = g.ferror_history(3)
    ao, ho = o.ferror_history(3)
    print(f"  gpu {ag} {hg:.6e}; oracle {ao} {ho:.6e}", flush=True)
    assert rel_err(ag, ao) < 1e-9
    assert abs(hg - ho) <= 1e-9 * abs(ho)
    assert ag[0] == g.scalars().ferror
    with pytest.raises(solver.PibError):
        g.ferror_history(4)
