"""Pinning the CPU oracle (no GPU): library routines, closed forms,
invariants and brute force, chosen so that a dropped term, a wrong sign or a
transposed index fails."""
import ctypes
import math

import numpy as np
import pytest
import scipy.linalg

from conftest import make_setup, rel_err
from pureibm_ps2d_fd1d_amd import _abi, solver


# ---------------------------------------------------------------- transforms
@pytest.mark.parametrize("my,mz", [(8, 8), (10, 6), (12, 20), (18, 10), (32, 24), (40, 30)])
def test_ff2rc_is_scaled_rfft2(orc, my, mz):
    a = np.random.default_rng(my + 7 * mz).standard_normal((my, mz))
    b = orc.ff2rc(a)
    ref = np.fft.rfft2(a.T).T / (my * mz)  # y (first, contiguous axis) is halved
    assert rel_err(b, ref) < 1e-14


@pytest.mark.parametrize("my,mz", [(8, 8), (10, 6), (12, 20), (18, 10), (32, 24)])
def test_ff2cr_row_column_convention(orc, my, mz):
    """Inverse c2c along z, then c2r along y ignoring Im of ky = 0, my/2:
    numpy's irfft2 follows the same row-column algorithm."""
    rng = np.random.default_rng(3 * my + mz)
    b = rng.standard_normal((my // 2 + 1, mz)) + 1j * rng.standard_normal((my // 2 + 1, mz))
    a = orc.ff2cr(b, my)
    ref = np.fft.irfft2(b.T, s=(mz, my)).T * (my * mz)
    assert rel_err(a, ref) < 1e-13
    # Im of the DC/Nyquist columns *after* the z pass is what gets dropped, so
    # changing Im(b[0, :]) does change the result (it mixes through z) ...
    b2 = b.copy()
    b2[0, :] += 1j * rng.standard_normal(mz)
    assert rel_err(orc.ff2cr(b2, my), a) > 1e-3
    # ... but a z-Hermitian-projected column has no visible imaginary part
    col = np.fft.fft(np.fft.ifft(b[0, :]).real)
    b3 = b.copy()
    b3[0, :] = col
    b4 = b3.copy()
    b4[0, :] = col + 1j * np.fft.fft(rng.standard_normal(mz))  # i * (z-Hermitian): imaginary after the z pass
    assert rel_err(orc.ff2cr(b4, my), orc.ff2cr(b3, my)) < 1e-13


def test_transform_round_trip_and_parseval(orc):
    rng = np.random.default_rng(5)
    a = rng.standard_normal((24, 16))
    b = orc.ff2rc(a)
    assert rel_err(orc.ff2cr(b, 24), a) < 1e-14
    # Parseval with the half-spectrum weights (ky = 0 and my/2 counted once)
    w = np.full(b.shape[0], 2.0)
    w[0] = w[-1] = 1.0
    e_spec = np.sum(w[:, None] * np.abs(b) ** 2) * a.size
    assert abs(e_spec - np.sum(a ** 2)) < 1e-12 * np.sum(a ** 2)


@pytest.mark.parametrize("n", [2, 3, 5, 8, 12, 30, 64, 96, 160])
def test_fft1d_matches_numpy(orc, n):
    x = np.random.default_rng(n).standard_normal(n) + 1j * np.random.default_rng(n + 1).standard_normal(n)
    assert rel_err(orc.fft1d(x, -1), np.fft.fft(x)) < 1e-14
    assert rel_err(orc.fft1d(x, +1), np.fft.ifft(x) * n) < 1e-14


# ------------------------------------------------------------- tridiagonal
def test_tridag_matches_banded_solver(orc):
    rng = np.random.default_rng(1)
    n = 50
    a, c = rng.uniform(0.2, 1, n), rng.uniform(0.2, 1, n)
    b = -2.5 - rng.uniform(0, 1, n)
    r = rng.standard_normal(n)
    x, st = orc.tridag(a, b, c, r)
    ab = np.zeros((3, n))
    ab[0, 1:] = c[:-1]
    ab[1] = b
    ab[2, :-1] = a[1:]
    assert st == 0
    assert rel_err(x, scipy.linalg.solve_banded((1, 1), ab, r)) < 1e-13


def test_cyclic_solve_matches_dense(orc):
    rng = np.random.default_rng(2)
    n = 33
    a, c = rng.uniform(0.2, 1, n), rng.uniform(0.2, 1, n)
    b = -2.5 - rng.uniform(0, 1, n)
    r = rng.standard_normal(n)
    x, st = orc.cyclic_solve(a, b, c, r)
    M = np.diag(b) + np.diag(a[1:], -1) + np.diag(c[:-1], 1)
    M[0, -1] = a[0]   # periodic corners, usteptime.F90:103-108
    M[-1, 0] = c[-1]
    assert st == 0
    assert rel_err(x, np.linalg.solve(M, r)) < 1e-12


def test_zero_pivot_reported(orc):
    a = np.ones(4)
    b = np.array([0.0, 1.0, 1.0, 1.0])
    x, st = orc.tridag(a, b, a, np.ones(4))
    assert st == _abi.PIB_ERR_PIVOT


# ------------------------------------------------------------- point sets
@pytest.mark.parametrize("d,count", [(12, 1158), (16, 2058), (20, 3190)])
def test_quadgener_counts_and_symmetry(orc, d, count):
    bnd, cd = orc.quadgener(d / 2.0)
    assert bnd.shape[1] == count
    assert np.allclose(np.linalg.norm(bnd, axis=0), 1.0, atol=2e-5)
    assert np.allclose(cd, bnd / np.linalg.norm(bnd, axis=0), atol=1e-15)
    pts = {tuple(np.round(p, 12)) for p in bnd.T}
    for s in [(-1, 1, 1), (1, -1, 1), (1, 1, -1)]:
        assert {tuple(np.round(np.array(p) * s, 12)) for p in pts} == pts


def test_host_quadgener_matches_oracle(orc):
    host = _abi.load_host()
    r = 8.0
    n = host.pib_host_quadgener(r, 1.6, 0, None, None)
    b = np.zeros((3, n), order="F")
    c = np.zeros((3, n), order="F")
    host.pib_host_quadgener(r, 1.6, n, _abi.dptr(b), _abi.dptr(c))
    bo, co = orc.quadgener(r)
    assert np.array_equal(b, bo) and np.array_equal(c, co)


# ------------------------------------------------------------- solid mask
def test_mask_brute_force(tmp_path, orc):
    st = make_setup(tmp_path, "MK", kind="random", my=24, dbydx=6, nbody=5, seed=8, re=20.0)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    m = o.field("mask")
    nx, my, mz = st.nx, st.my, st.mz
    I, J, K = np.meshgrid(np.arange(1, nx + 1), np.arange(1, my + 1), np.arange(1, mz + 1), indexing="ij")
    solid = np.zeros_like(I, dtype=bool)
    L = np.array([nx, my, mz])
    for p, r in zip(st.xc, st.radbdy):
        d = np.stack([I - p[0], J - p[1], K - p[2]])
        d -= (L[:, None, None, None] * np.round(d / L[:, None, None, None]))
        solid |= np.sqrt((d ** 2).sum(0)) - r <= 1e-35
    assert np.array_equal(m == 0.0, solid)
    assert o.count_solid == solid.sum()


# ------------------------------------------------------------- operators
def _bare_oracle(orc, nx, my, mz, dx=0.1, vis=0.05, dt=0.01, scheme=0, limiter=0):
    lib = orc.load()
    c = _abi.PibConfig()
    c.nx, c.my, c.mz = nx, my, mz
    c.dx, c.vis, c.cfl_max = dx, vis, 0.5
    c.discr_scheme, c.limiter = scheme, limiter
    c.revernorm, c.dorpr, c.soft_start = 1, 1, 1
    c.ucharmod, c.char_length, c.dia, c.dr, c.voldom = 1.0, 1.0, 1.0, 1.0, 1.0
    ctx = ctypes.c_void_p()
    assert lib.orc_create(ctypes.byref(c), ctypes.byref(ctx)) == 0
    assert lib.orc_set_particles(ctx, 0, None, None, None, None, None, None) == 0
    s = _abi.PibScalars()
    s.dt, s.mean_volfrac = dt, 1.0
    lib.orc_set_scalars(ctx, ctypes.byref(s))
    return lib, ctx


def _put(lib, ctx, fid, arr):
    a = np.asfortranarray(arr.astype(np.complex128))
    assert lib.orc_set_field(ctx, fid, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0


def _get(lib, ctx, fid, shape, cplx=True):
    a = np.zeros(shape, dtype=np.complex128 if cplx else np.float64, order="F")
    assert lib.orc_get_field(ctx, fid, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    return a


def test_helmholtz_symbol(orc):
    """A divergence-free single mode u3 ~ e^{i theta i} (kz = 0) is advanced by
    CN diffusion alone: u_new = u (1 + dt nu a lam)/(1 - dt nu b lam) with the
    discrete symbol lam = -4 sin^2(theta/2)/dx^2 - ky^2."""
    nx, my, mz, dx, vis, dt = 16, 8, 8, 0.1, 0.05, 0.01
    lib, ctx = _bare_oracle(orc, nx, my, mz, dx, vis, dt)
    m, jy = 3, 2
    th = 2 * math.pi * m / nx
    u = np.zeros((nx + 1, my // 2 + 1, mz, 3), dtype=np.complex128)
    ii = np.arange(nx + 1)
    u[:, jy, 0, 2] = np.exp(1j * th * ii)
    _put(lib, ctx, 0, u)
    assert lib.orc_velstep(ctx, 1) == 0
    un = _get(lib, ctx, 0, u.shape)
    ky = 2 * math.pi * jy / (my * dx)
    lam = -4 * math.sin(th / 2) ** 2 / dx ** 2 - ky ** 2
    g = (1 + dt * vis * 0.5 * lam) / (1 - dt * vis * 0.5 * lam)
    assert rel_err(un[:, jy, 0, 2], g * u[:, jy, 0, 2]) < 1e-12
    assert np.max(np.abs(un[:, :, :, :2])) < 1e-14
    lib.orc_destroy(ctx)


def test_projection_leaves_divergence_free(orc):
    """After velstep the staggered divergence is round-off for any real input
    field (spectra of real fields: the mean mode is real)."""
    nx, my, mz = 12, 8, 10
    lib, ctx = _bare_oracle(orc, nx, my, mz)
    rng = np.random.default_rng(4)
    shp = (nx + 1, my // 2 + 1, mz, 3)
    for fid in (0, 2, 4):  # u, nl, ff
        f = np.zeros(shp, dtype=np.complex128)
        for i in range(nx):
            for c in range(3):
                f[i, :, :, c] = orc.ff2rc(rng.standard_normal((my, mz)))
        f[nx] = f[0]
        _put(lib, ctx, fid, f)
    assert lib.orc_velstep(ctx, 1) == 0
    s = _abi.PibScalars()
    lib.orc_get_scalars(ctx, ctypes.byref(s))
    u = _get(lib, ctx, 0, shp)
    assert s.divmax < 1e-10 * np.max(np.abs(u))
    lib.orc_destroy(ctx)


def test_form_nl_center_manufactured(orc):
    """u = U0 + a cos(2 pi m x/Lx), v = b sin(2 pi y/Ly), w = 0:
    nl_1 = -[(u^2(i+1) - u^2(i-1))/(2dx) + u dv/dy] with exact spectral
    y-derivative and the discrete central x-difference."""
    nx, my, mz, dx = 12, 8, 8, 0.25
    lib, ctx = _bare_oracle(orc, nx, my, mz, dx=dx)
    Lx, Ly = nx * dx, my * dx
    x = (np.arange(1, nx + 1) - 1) * dx
    y = (np.arange(1, my + 1) - 1) * dx
    a, b, m, U0 = 0.3, 0.2, 2, 0.7
    uphys = np.zeros((nx, my, mz, 3))
    uphys[..., 0] = (a * np.cos(2 * np.pi * m * x / Lx))[:, None, None]
    uphys[..., 1] = (b * np.sin(2 * np.pi * y / Ly))[None, :, None]
    uspec = np.zeros((nx + 1, my // 2 + 1, mz, 3), dtype=np.complex128)
    for i in range(nx):
        for c in range(3):
            uspec[i, :, :, c] = orc.ff2rc(uphys[i, :, :, c])
    uspec[nx] = uspec[0]
    _put(lib, ctx, 0, uspec)
    s = _abi.PibScalars()
    lib.orc_get_scalars(ctx, ctypes.byref(s))
    s.umean[0] = U0
    lib.orc_set_scalars(ctx, ctypes.byref(s))
    assert lib.orc_compute_new_timestep(ctx, 1) == 0
    assert lib.orc_nonlinear(ctx, 1) == 0
    nlb = _get(lib, ctx, 7, (nx + 1, my, mz, 3), cplx=False)
    u = U0 + uphys[..., 0]
    v = uphys[..., 1]
    dudx2 = (np.roll(u ** 2, -1, 0) - np.roll(u ** 2, 1, 0)) / (2 * dx)
    dvdy = (b * 2 * np.pi / Ly * np.cos(2 * np.pi * y / Ly))[None, :, None]
    ref1 = -(dudx2 + u * dvdy)
    assert rel_err(nlb[:nx, :, :, 0], ref1) < 1e-12
    # nl_2 = -(d(uv)/dx + d(vv)/dy): central x-difference of u v, exact y-derivative of v^2
    duvdx = (np.roll(u * v, -1, 0) - np.roll(u * v, 1, 0)) / (2 * dx)
    dv2dy = (2 * b * b * 2 * np.pi / Ly * np.sin(2 * np.pi * y / Ly) * np.cos(2 * np.pi * y / Ly))[None, :, None]
    assert rel_err(nlb[:nx, :, :, 1], -(duvdx + dv2dy)) < 1e-12
    lib.orc_destroy(ctx)


def test_forcing_redistribution_conserves(tmp_path, orc):
    """Every reversal force lands on the grid (solid share + redistributed
    fluid share), so fr - frmean sums to zero over the box and the plane means
    of ff (its (0,0) modes) sum to zero along x."""
    st = make_setup(tmp_path, "CS", kind="random", my=24, dbydx=6, nbody=3, seed=5, re=30.0)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    o.step(3)
    fr = o.field("fr")[: st.nx]
    s = o.scalars()
    assert abs(fr[..., 0].sum()) < 1e-10 * np.abs(fr[..., 0]).sum()
    ff00 = o.field("ff")[: st.nx, 0, 0, :]
    assert np.max(np.abs(ff00.sum(0))) < 1e-12 * np.max(np.abs(fr))
    assert abs(s.frmean[0]) > 0


@pytest.mark.parametrize("limiter", [0, 1])
def test_upwind_flux_is_fromm_for_constant_advection(orc, limiter):
    """u = U > 0 constant, v = A cos(2 pi m x/Lx), w = 0.  The Rusanov flux of
    the MUSCL states reduces to pure upwinding, F(i+1/2) = U (v_i + s_i dx/2):
    without a limiter s_i is the central slope (Fromm's scheme); with minmod
    it is the central slope times phi(r) = max(0, min(1, r)),
    r = (v_i - v_{i-1})/(v_{i+1} - v_i).  nl_2 = -(F(i+1/2) - F(i-1/2))/dx,
    and nl_1 = 0 (u is constant, v does not depend on y)."""
    nx, my, mz, dx = 16, 8, 8, 0.25
    lib, ctx = _bare_oracle(orc, nx, my, mz, dx=dx, scheme=1, limiter=limiter)
    U, A, m = 0.8, 0.3, 1
    x = np.arange(nx) * dx
    v = A * np.cos(2 * np.pi * m * x / (nx * dx))
    uspec = np.zeros((nx + 1, my // 2 + 1, mz, 3), dtype=np.complex128)
    uspec[:nx, 0, 0, 1] = v
    uspec[nx] = uspec[0]
    _put(lib, ctx, 0, uspec)
    s = _abi.PibScalars()
    lib.orc_get_scalars(ctx, ctypes.byref(s))
    s.umean[0] = U
    lib.orc_set_scalars(ctx, ctypes.byref(s))
    assert lib.orc_compute_new_timestep(ctx, 1) == 0
    assert lib.orc_nonlinear(ctx, 1) == 0
    nlb = _get(lib, ctx, 7, (nx + 1, my, mz, 3), cplx=False)
    vm, vp = np.roll(v, 1), np.roll(v, -1)
    slope = (vp - vm) / (2 * dx)
    if limiter:
        with np.errstate(divide="ignore", invalid="ignore"):
            r = (v - vm) / (vp - v)
        slope = slope * np.clip(r, 0.0, 1.0)
    F = U * (v + slope * dx / 2)  # at i+1/2
    ref = -(F - np.roll(F, 1)) / dx
    for j in range(my):
        for k in range(mz):
            assert rel_err(nlb[:nx, j, k, 1], ref) < 1e-12
    assert np.max(np.abs(nlb[:nx, :, :, 0])) < 1e-14 * U * A / dx
    lib.orc_destroy(ctx)


def _one_sphere(tmp_path, orc, name):
    st = make_setup(tmp_path, name, kind="simple", my=32, dbydx=8, re=20.0)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    m = 0
    r = st.radbdy[m]
    xl = st.xc[m][:, None] + st.bndpts * r  # (3, nbnd) grid units, node i at i
    da1 = 4 * np.pi * (r * st.config.dx) ** 2 / st.nbnd
    return st, o, xl, da1


def test_pressure_drag_of_a_linear_pressure_field(tmp_path, orc):
    """With u = 0 and pbcp = G * x on the pressure grid (node i at x = (i+1/2) dx)
    the trilinear interpolation is exact, so pres = -sum_l G x_l cd_l dA and,
    as the point set refines, -G * volume in x."""
    st, o, xl, da1 = _one_sphere(tmp_path, orc, "PD")
    G, dx = 0.7, st.config.dx
    o.compute_new_timestep(1)
    i = np.arange(1, st.nx + 1)
    pb = np.broadcast_to((G * (i + 0.5) * dx)[:, None, None], (st.nx, st.my, st.mz))
    o.set_field("pbcp", pb)
    o.nonlinear(1)
    pres = o.forces()["pres"][0]
    ref = -np.sum(G * xl[0] * dx * st.cd * da1, axis=1)
    assert rel_err(pres, ref) < 1e-12
    vol = 4.0 / 3.0 * np.pi * (st.radbdy[0] * dx) ** 3
    assert abs(pres[0] + G * vol) < 0.05 * G * vol
    assert np.max(np.abs(o.forces()["visc"])) < 1e-14 * G * vol


def test_viscous_drag_of_a_shear_mode(tmp_path, orc):
    """u_x = A cos(k y) gives omega = (0, 0, A k sin(k y)) exactly at the nodes
    (spectral y-derivative of one mode); the viscous traction on a no-slip
    surface is nu * (omega x n) dA (n the outward normal), with omega
    interpolated trilinearly (boundary_condition.F90:1579-1613)."""
    st, o, xl, da1 = _one_sphere(tmp_path, orc, "VD")
    A, dx, my, mz = 0.2, st.config.dx, st.my, st.mz
    k = 2 * np.pi / (my * dx)
    yj = np.arange(my) * dx
    plane = np.broadcast_to((A * np.cos(k * yj))[:, None], (my, mz))
    b = np.fft.rfft2(plane.T).T / (my * mz)
    u = np.zeros((st.nx + 1, my // 2 + 1, mz, 3), dtype=np.complex128)
    u[:, :, :, 0] = b[None]
    o.set_field("u", u)
    o.compute_new_timestep(1)
    o.nonlinear(1)
    visc = o.forces()["visc"][0]
    # trilinear along y (omega is constant in x and z): node j (1-based) at y = j
    y = xl[1]
    jb = np.floor(y).astype(int)
    eta = y - jb
    wz_node = lambda j: A * k * np.sin(k * ((j - 1) % my) * dx)
    wz = (1 - eta) * wz_node(jb) + eta * wz_node(jb + 1)
    cd = st.cd
    df = np.stack([-wz * cd[1], wz * cd[0], np.zeros_like(wz)])  # omega x n
    ref = np.sum(st.config.vis * df * da1, axis=1)
    assert abs(ref[0]) > 1e-3 * st.config.vis * A * k * np.sum(da1 * np.abs(cd[1]))
    assert rel_err(visc, ref) < 1e-12
    assert np.max(np.abs(o.forces()["pres"])) < 1e-14 * np.max(np.abs(ref))


def test_compute_new_timestep_masked_means(tmp_path, orc):
    """Brute-force fluid/solid/box means and max|u| of ubcp over the mask."""
    st = make_setup(tmp_path, "CN", kind="random", my=24, dbydx=6, nbody=4, seed=12, re=20.0)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    rng = np.random.default_rng(3)
    nx, my, mz = st.nx, st.my, st.mz
    u = np.zeros((nx + 1, my // 2 + 1, mz, 3), dtype=np.complex128)
    for i in range(nx):
        for c in range(3):
            u[i, :, :, c] = np.fft.rfft2(rng.standard_normal((mz, my))).T / (my * mz)
    u[nx] = u[0]
    o.set_field("u", u)
    s0 = o.scalars()
    o.compute_new_timestep(1)
    s = o.scalars()
    ub = o.field("ubcp")[:nx]
    fl = o.field("mask") != 0
    for c in range(3):
        assert abs(s.ufmean[c] - ub[..., c][fl].sum() / fl.sum()) < 1e-13
        assert abs(s.usmean_temp[c] - ub[..., c][~fl].sum() / (~fl).sum()) < 1e-13
        assert abs(s.umean_temp[c] - ub[..., c].mean()) < 1e-13
    assert abs(s.umax_tmp - np.abs(ub).max()) < 1e-14
    assert s.ufmean_old[0] == s0.ufmean[0]


@pytest.mark.parametrize("move", [0, 1])
def test_new_timestep_moving_particle_cfl(tmp_path, orc, move):
    """dependent_functions.F90:1150-1165: dtnew = cfl_max*dx/umax_tmp when the
    fluid breaks the CFL bound; with moving particles the same formula (still
    from the fluid's umax_tmp) also applies when the fastest particle breaks
    it, even where the fluid alone keeps dt.  Case: up_max*dt/dx > cfl_max >=
    umax_tmp*dt/dx, so the new dt differs between the two branches."""
    st = make_setup(tmp_path, f"MV{move}", kind="random", my=24, dbydx=6, nbody=4, seed=12, re=20.0,
                    adaptive=True)
    st.config.move_particles = move
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    s0 = o.scalars()
    cfl_max, dx = st.config.cfl_max, st.config.dx
    up = 4.0 * cfl_max * dx / s0.dt  # the fastest particle moves 4 cfl_max cells a step
    vel = np.zeros((st.nbody, 3))
    vel[2] = [0.6 * up, 0.0, 0.8 * up]  # |v| = up, composed over components
    vel[0] = [0.1 * up, 0.1 * up, 0.0]
    o.update_particles(st.xc, vel)
    o.compute_new_timestep(1)
    s = o.scalars()
    assert s.umax_tmp * s0.dt / dx <= cfl_max  # the fluid alone keeps dt
    if move:
        assert s.dt == cfl_max * dx / s.umax_tmp and s.dt != s0.dt
    else:
        assert s.dt == s0.dt
    assert s.dtnew == s.dt
    o.close()


@pytest.mark.parametrize("scheme", ["center", "upwind"])
def test_threads_do_not_change_a_bit(tmp_path, orc, scheme):
    """The oracle's plane transforms and mode solves run under OpenMP; the
    sums stay serial, so one thread and four give the same bits."""
    lib = orc.load()
    kw = dict(kind="random", my=24, dbydx=6, nbody=3, seed=8, re=40.0, discr_scheme=scheme)
    out = []
    nt0 = lib.orc_get_threads()
    try:
        for nt in (1, 4):
            lib.orc_set_threads(nt)
            st = make_setup(tmp_path, f"TH{nt}", **kw)
            o = solver.Solver(st, lib=lib, prefix="orc_")
            o.step(3)
            out.append((o.field("u"), o.field("p"), o.forces()["force"], o.scalars().norm_drag))
            o.close()
    finally:
        lib.orc_set_threads(nt0)
    for a, b in zip(*out):
        assert np.array_equal(a, b)
