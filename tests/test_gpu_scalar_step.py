"""The whole scalar step (pib_scalstep: nlphi_graduphi with form_nlphi and
compute_phim_heat_ratio, bcsetscal with calc_diffn, calc_bndrev_data and
calc_inner_rev_data, then the mode solve; stepthescalar.F90:51-95) on the GPU
against the oracle (pinned in test_scalar_step_oracle.py), through the C-ABI.

Each case couples the scalar to the flow as the reference's time loop does
(dtibm_main_scal_flo.F90:426-515): per stage compute_new_timestep,
nonlinear, velstep, then the scalar step on the stage's ubcp.  The scalar
starts from phi = 1 plus a smooth random perturbation, the twisted wrap
phif(nx+1) = phif(1)/heat_ratio holding for the starting ratio.  Cases:
central and upwind x-fluxes, reversal pairs and/or surface forcing,
include_frmeanfluid, AB2 and RK3, compiled and generic plane shapes and the
HBM plane form, beds with spheres across the x wrap.

Tolerances, FP64: relative 1e-10 to the field maximum on phif, nlphif,
onlphif and ffphi, 1e-12 on the heat ratios, 1e-9 on the surface fluxes
(the transforms and the spreading sum in orders of their own).  FP32-bulk
against FP64 on the GPU: 1e-4 relative on phif and the mean surface flux.
"""
import numpy as np
import pytest

from conftest import make_setup, rel_err
from pureibm_ps2d_fd1d_amd import _abi, solver

pytestmark = pytest.mark.gpu

GAMMA = 0.012 / 0.7  # vis / Pr_or_Sc_nu (scalar_init.F90:73), the MIS deck's Pr


def _phi0(st, hr, seed):
    """phif of phi = 1 + 0.05 sum of a few smooth random modes; plane nx+1 =
    plane 1 / hr."""
    nx, my, mz = st.nx, st.my, st.mz
    rng = np.random.default_rng(seed)
    x = np.arange(nx)[:, None, None] / nx
    y = np.arange(my)[None, :, None] / my
    z = np.arange(mz)[None, None, :] / mz
    ph = np.ones((nx, my, mz))
    for _ in range(4):
        kx, ky, kz = rng.integers(0, 3, 3)
        ph += 0.05 * rng.standard_normal() * np.cos(2 * np.pi * (kx * x + ky * y + kz * z) + rng.uniform(0, 6.3))
    spec = np.stack([np.fft.rfft2(ph[i].T).T / (my * mz) for i in range(nx)])
    return np.concatenate([spec, spec[:1] / hr])


def _run_pair(st, orc, nsteps, hr=1.0, seed=1, tune_fft=None):
    g = solver.Solver(st)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    if tune_fft is not None:
        assert g.lib.pib_set_tuning(g.ctx, _abi.PIB_TUNE_FFT, tune_fft) == 0
    phi = _phi0(st, hr, seed)
    for b in (g, o):
        b.scalar_set(GAMMA, hr, hr)
        b.scalar_surface(np.zeros(st.nbody))
        b.set_field("phi", phi)
    itr = 3 if st.config.rk3 else 1
    for n in range(nsteps):
        for rks in range(1, itr + 1):
            for b in (g, o):
                b.compute_new_timestep(rks)
                b.nonlinear(rks)
                b.velstep(rks)
                b.scalstep(rks)
        _compare(g, o, f"step {n + 1}")
    return g, o


def _compare(g, o, tag):
    errs = {f: rel_err(g.field(f), o.field(f)) for f in ("phi", "nlphi", "onlphi", "ffphi")}
    rg, ro = g.scalar_get(), o.scalar_get()
    fg, fo = g.scalar_flux(), o.scalar_flux()
    efb = rel_err(fg[0], fo[0]) if len(fo[0]) else 0.0
    print(f"  {tag}: " + " ".join(f"{k} {v:.1e}" for k, v in errs.items()) +
          f" ratios {rg[0]:.12f}/{ro[0]:.12f} flux {efb:.1e} fluid mean {fg[2]:.12f}", flush=True)
    for k, v in errs.items():
        assert v < 1e-10, (tag, k)
    for a, b in zip(rg[:3], ro[:3]):
        assert abs(a - b) <= 1e-12 * abs(b), tag
    assert efb < 1e-9, tag
    assert abs(fg[1] - fo[1]) <= 1e-9 * max(abs(fo[1]), 1e-30), tag
    assert abs(fg[2] - fo[2]) <= 1e-12 * abs(fo[2]) and abs(fg[3] - fo[3]) <= 1e-12 * max(abs(fo[3]), 1.0), tag


CASES = {
    "center_rpr_32": dict(my=32, dbydx=8, nbody=3, seed=41, re=20.0),
    "upwind_rpr_24": dict(my=24, dbydx=6, nbody=4, seed=42, re=20.0, discr_scheme="upwind"),
    "upwind_rpr_bnd_rk3_32": dict(my=32, dbydx=8, nbody=3, seed=43, re=20.0, discr_scheme="upwind", rk3=True,
                                  dobnd=True),
    "center_bnd_frmeanfluid_24": dict(my=24, dbydx=6, nbody=3, seed=44, re=20.0, dorpr=False, dobnd=True,
                                      include_frmeanfluid=True),
    "upwind_rpr_64": dict(my=64, dbydx=12, nbody=8, seed=45, re=30.0, discr_scheme="upwind"),
}


@pytest.mark.parametrize("case", list(CASES))
def test_scalstep_matches_the_oracle(tmp_path, orc, case):
    kw = dict(CASES[case])
    st = make_setup(tmp_path, case.upper(), kind="random", **kw)
    print(f"scalstep {case}: {st.nx}^3, {st.nbody} spheres", flush=True)
    g, o = _run_pair(st, orc, 3, hr=1.0 if "center" in case else 1.03, seed=len(case))
    g.close()
    o.close()


def test_scalstep_with_planes_in_hbm(tmp_path, orc):
    st = make_setup(tmp_path, "HBM", kind="random", my=32, dbydx=8, nbody=3, seed=46, re=20.0,
                    discr_scheme="upwind")
    print("scalstep 32^3 with the plane transforms in HBM slots", flush=True)
    g, o = _run_pair(st, orc, 2, hr=1.02, tune_fft=2)
    g.close()
    o.close()


def test_fp32_bulk_scalar_against_fp64(tmp_path):
    res = {}
    for prec in (_abi.PIB_FP64, _abi.PIB_FP32_BULK):
        st = make_setup(tmp_path / str(prec), "F32", kind="random", precision=prec, my=64, dbydx=12, nbody=8,
                        seed=47, re=30.0, discr_scheme="upwind")
        g = solver.Solver(st)
        g.scalar_set(GAMMA, 1.0, 1.0)
        g.scalar_surface(np.zeros(st.nbody))
        g.set_field("phi", _phi0(st, 1.0, 5))
        for _ in range(10):
            g.compute_new_timestep(1)
            g.nonlinear(1)
            g.velstep(1)
            g.scalstep(1)
        res[prec] = (g.field("phi"), g.scalar_flux(), g.scalar_get())
        g.close()
    e = rel_err(res[1][0], res[0][0])
    ef = abs(res[1][1][1] - res[0][1][1]) / abs(res[0][1][1])
    print(f"FP32-bulk vs FP64 after 10 steps: phi {e:.1e}, flux_global {ef:.1e}, "
          f"heat ratio {res[1][2][0]:.9f} / {res[0][2][0]:.9f}", flush=True)
    assert e < 1e-4 and ef < 1e-4
