"""Closed-form pins of the oracle's bcset and velstep pieces (no GPU).

The GPU tests compare the kernels with the oracle routine by routine, so a
mistake shared by both would pass them.  Each test here fixes an oracle
routine by what the mathematics gives for a manufactured input: the discrete
symbol of a single Fourier mode, an exact derivative of a band-limited field,
the reflection of a uniform flow at a sphere, the controller's response with
no surface drag.  A dropped term, a wrong sign, a wrong average or a
transposed index changes the expected value.
"""
import ctypes
import math

import numpy as np
import pytest

from conftest import make_setup, rel_err
from pureibm_ps2d_fd1d_amd import _abi, solver

AB2 = (0.5, 0.5, 1.5, -0.5)  # coef(1, 1:4) of the AB2/CN step, initialize_flo.F90:861-885


def _bare(orc, nx, my, mz, dx=0.1, vis=0.05, dt=0.01):
    lib = orc.load()
    c = _abi.PibConfig()
    c.nx, c.my, c.mz = nx, my, mz
    c.dx, c.vis, c.cfl_max = dx, vis, 0.5
    c.revernorm, c.dorpr, c.soft_start = 1, 1, 1
    c.ucharmod, c.char_length, c.dia, c.dr, c.voldom = 1.0, 1.0, 1.0, 1.0, 1.0
    ctx = ctypes.c_void_p()
    assert lib.orc_create(ctypes.byref(c), ctypes.byref(ctx)) == 0
    assert lib.orc_set_particles(ctx, 0, None, None, None, None, None, None) == 0
    s = _abi.PibScalars()
    s.dt, s.mean_volfrac = dt, 1.0
    lib.orc_set_scalars(ctx, ctypes.byref(s))
    return lib, ctx


def _put(lib, ctx, name, arr):
    a = np.asfortranarray(arr.astype(np.complex128))
    assert lib.orc_set_field(ctx, _abi.FIELDS[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0


def _get(lib, ctx, name, shape):
    a = np.zeros(shape, dtype=np.complex128, order="F")
    assert lib.orc_get_field(ctx, _abi.FIELDS[name], a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    return a


def _wavenumbers(my, mz, dx):
    ky = 2 * np.pi * np.arange(my // 2 + 1) / (my * dx)
    kz = 2 * np.pi * np.fft.fftfreq(mz, 1.0 / mz) / (mz * dx)
    return ky, kz


def _r2c(plane):
    """ff2rc of a (my, mz) plane: numpy's rfft2 on the transposed axes, scaled."""
    my, mz = plane.shape
    return np.fft.rfft2(plane.T).T / (my * mz)


def _c2r(spec, my):
    mz = spec.shape[1]
    return np.fft.irfft2(spec.T, s=(mz, my)).T * (my * mz)


# ------------------------------------------------------------ velstep
def test_pressure_predictor_symbol(orc):
    """poisson.F90:55-239.  With forcing terms e^{i theta i} on the velocity
    nodes the right-hand side at the pressure point i+1/2 is the staggered
    divergence [c3 div nl + div ff + c4 div onl]/(c1+c2):
        x: (F(i+1) - F(i))/dx,   y, z: i k (F(i) + F(i+1))/2,
    and the Poisson matrix (a = c = 1/dx^2 - w2/4, b = -2/dx^2 - w2/2) has the
    symbol -4 sin^2(theta/2)/dx^2 - w2 cos^2(theta/2).  The (0,0) mode is
    solved on rows 1..nx-1 with p(nx) = 0 and a real result: the periodic
    solution shifted by its value at nx."""
    nx, my, mz, dx = 16, 8, 8, 0.1
    lib, ctx = _bare(orc, nx, my, mz, dx)
    c1, c2, c3, c4 = AB2
    ky, kz = _wavenumbers(my, mz, dx)
    m, jy, jz = 3, 2, 7  # kz index 7 of 8 is the negative wavenumber -1
    th = 2 * np.pi * m / nx
    ii = np.arange(nx + 1)
    e = np.exp(1j * th * ii)
    shp = (nx + 1, my // 2 + 1, mz, 3)
    rng = np.random.default_rng(11)
    amp = {f: rng.standard_normal(3) + 1j * rng.standard_normal(3) for f in ("nl", "ff", "onl")}
    fields = {}
    for f in ("nl", "ff", "onl"):
        a = np.zeros(shp, dtype=np.complex128)
        for c in range(3):
            a[:, jy, jz, c] = amp[f][c] * e
        fields[f] = a
    # (0,0) mode: a real cos wave in nl_x only
    m0, a0 = 2, 0.8
    th0 = 2 * np.pi * m0 / nx
    fields["nl"][:, 0, 0, 0] = a0 * np.cos(th0 * ii)
    for f, a in fields.items():
        _put(lib, ctx, f, a)
    assert lib.orc_pressure_modes(ctx, 1) == 0
    p = _get(lib, ctx, "p", (nx, my // 2 + 1, mz))

    kyv, kzv = ky[jy], kz[jz]
    w2 = kyv ** 2 + kzv ** 2

    def div(F):
        return F[0] * (np.exp(1j * th) - 1) / dx + 1j * (kyv * F[1] + kzv * F[2]) * (1 + np.exp(1j * th)) / 2

    rhs = (c3 * div(amp["nl"]) + div(amp["ff"]) + c4 * div(amp["onl"])) / (c1 + c2)
    lam = -4 * np.sin(th / 2) ** 2 / dx ** 2 - w2 * np.cos(th / 2) ** 2
    ref = rhs / lam * e[:nx]
    assert rel_err(p[:, jy, jz], ref) < 1e-13
    # the (0,0) mode
    rhs0 = c3 * a0 * (np.cos(th0 * (ii[:nx] + 1)) - np.cos(th0 * ii[:nx])) / dx / (c1 + c2)
    per = rhs0 / (-4 * np.sin(th0 / 2) ** 2 / dx ** 2)
    ref0 = per - per[nx - 1]
    assert rel_err(p[:, 0, 0].real, ref0) < 1e-12
    assert np.all(p[:, 0, 0].imag == 0.0)
    # no other mode is touched
    mask = np.ones(p.shape, bool)
    mask[:, jy, jz] = mask[:, 0, 0] = False
    assert np.all(p[mask] == 0.0)
    lib.orc_destroy(ctx)


def test_gradient_forcing_is_absorbed_by_the_pressure(orc):
    """A forcing ff = G q that is the discrete gradient of a pressure-point
    field q (x: (q(i) - q(i-1))/dx at node i, y and z: i k (q(i-1) + q(i))/2)
    is balanced by the pressure alone: after the whole velstep u stays zero
    and p = q/(c1+c2) (for the (0,0) mode shifted so that p(nx) = 0).  This
    fixes ustep's pressure gradient and divcorr's update against the
    predictor's divergence and symbol."""
    nx, my, mz, dx, dt = 12, 8, 6, 0.1, 0.02
    lib, ctx = _bare(orc, nx, my, mz, dx, dt=dt)
    c1, c2 = AB2[:2]
    ky, kz = _wavenumbers(my, mz, dx)
    shp = (nx + 1, my // 2 + 1, mz, 3)
    ff = np.zeros(shp, dtype=np.complex128)
    ii = np.arange(nx)
    qs = {}
    for (jy, jz, m, Q) in ((1, 2, 2, 0.6 - 0.3j), (3, 0, 5, -0.2 + 0.9j), (0, 0, 1, 0.5)):
        q = Q * np.exp(2j * np.pi * m * ii / nx)
        if jy == jz == 0:
            q = q.real.astype(np.complex128)
        qs[(jy, jz)] = q
        qm = np.roll(q, 1)  # q(i-1), periodic
        ff[:nx, jy, jz, 0] = (q - qm) / dx
        ff[:nx, jy, jz, 1] = 1j * ky[jy] * (qm + q) / 2
        ff[:nx, jy, jz, 2] = 1j * kz[jz] * (qm + q) / 2
    ff[nx] = ff[0]
    _put(lib, ctx, "ff", ff)
    assert lib.orc_velstep(ctx, 1) == 0
    u = _get(lib, ctx, "u", shp)
    p = _get(lib, ctx, "p", (nx, my // 2 + 1, mz))
    assert np.max(np.abs(u)) < 1e-13 * dt * np.max(np.abs(ff))
    for (jy, jz), q in qs.items():
        ref = q / (c1 + c2)
        if jy == jz == 0:
            ref = ref - ref[nx - 1]
        assert rel_err(p[:, jy, jz], ref) < 1e-12
    lib.orc_destroy(ctx)


def test_velstep_removes_a_gradient_velocity(orc):
    """A velocity u0 = G psi that is a discrete gradient (no forcing) has no
    predicted pressure; CN diffusion takes it to u* = (E/H) u0 per mode, with
    E = 1 - dt nu c1 Lam, H = 1 + dt nu c2 Lam and Lam = 4 sin^2(theta/2)/dx^2
    + w2; divcorr's projection removes it entirely, phi = (E/H) psi/(dt(c1+c2)),
    and the pressure update p += phi - dt c2 nu (delta^2 phi - w2 phi) = H phi
    leaves p = E psi/(dt(c1+c2)) (usteptime.F90:82-188, poisson.F90:241-438)."""
    nx, my, mz, dx, dt, vis = 12, 8, 6, 0.1, 0.02, 0.05
    lib, ctx = _bare(orc, nx, my, mz, dx, vis=vis, dt=dt)
    c1, c2 = AB2[:2]
    ky, kz = _wavenumbers(my, mz, dx)
    shp = (nx + 1, my // 2 + 1, mz, 3)
    u = np.zeros(shp, dtype=np.complex128)
    ii = np.arange(nx)
    cases = {}
    for (jy, jz, m, Q) in ((1, 2, 2, 0.6 - 0.3j), (2, 5, 1, 0.1 + 0.4j), (0, 0, 3, 0.5)):
        psi = Q * np.exp(2j * np.pi * m * ii / nx)
        if jy == jz == 0:
            psi = psi.real.astype(np.complex128)
        pm = np.roll(psi, 1)
        u[:nx, jy, jz, 0] = (psi - pm) / dx
        u[:nx, jy, jz, 1] = 1j * ky[jy] * (pm + psi) / 2
        u[:nx, jy, jz, 2] = 1j * kz[jz] * (pm + psi) / 2
        lam = 4 * np.sin(np.pi * m / nx) ** 2 / dx ** 2 + ky[jy] ** 2 + kz[jz] ** 2
        cases[(jy, jz)] = (psi, 1 - dt * vis * c1 * lam, 1 + dt * vis * c2 * lam)
    u[nx] = u[0]
    _put(lib, ctx, "u", u)
    assert lib.orc_velstep(ctx, 1) == 0
    un = _get(lib, ctx, "u", shp)
    p = _get(lib, ctx, "p", (nx, my // 2 + 1, mz))
    assert np.max(np.abs(un)) < 1e-12 * np.max(np.abs(u))
    for (jy, jz), (psi, E, H) in cases.items():
        ref = E * psi / (dt * (c1 + c2))
        if jy == jz == 0:  # phi is pinned to zero at nx; H of a constant is 1
            ref = ref - (E / H) * psi[nx - 1] / (dt * (c1 + c2))
        assert rel_err(p[:, jy, jz], ref) < 1e-12, (jy, jz)
    lib.orc_destroy(ctx)


# ------------------------------------------------------- bcset fields
def _sphere(tmp_path, orc, name, **kw):
    kw.setdefault("re", 20.0)
    st = make_setup(tmp_path, name, kind="simple", my=32, dbydx=8, **kw)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    return st, o


def _spec_of(phys, nplanes):
    """Spectral (nplanes, my2, mz) field of physical (nplanes, my, mz) planes."""
    return np.stack([_r2c(phys[i]) for i in range(nplanes)])


def test_calc_visc_symbol(tmp_path, orc):
    """boundary_condition.F90:1767-1815: diffn = nu (delta_x^2 u - w2 u),
    periodic in x (plane nx+1 takes plane 2 as its right neighbour).  For
    u ~ e^{i theta i} at one (ky, kz) mode, diffn = -nu (4 sin^2(theta/2)/dx^2
    + w2) times the physical field of u."""
    st, o = _sphere(tmp_path, orc, "VS")
    nx, my, mz, dx, vis = st.nx, st.my, st.mz, st.config.dx, st.config.vis
    ky, kz = _wavenumbers(my, mz, dx)
    shp = (nx + 1, my // 2 + 1, mz, 3)
    u = np.zeros(shp, dtype=np.complex128)
    ii = np.arange(nx + 1)
    modes = [((2, 3, 4), 0.3 + 0.1j, 0), ((1, 0, 31), -0.2j, 1), ((0, 5, 2), 0.25, 2)]
    for (m, jy, jz), A, c in modes:
        u[:, jy, jz, c] = A * np.exp(2j * np.pi * m * ii / nx)
    o.set_field("u", u)
    o.compute_new_timestep(1)
    o.nonlinear(1)
    diffn = o.field("diffn")
    for (m, jy, jz), A, c in modes:
        lam = -(4 * np.sin(np.pi * m / nx) ** 2 / dx ** 2 + ky[jy] ** 2 + kz[jz] ** 2)
        ref = np.stack([vis * lam * _c2r(u[i, :, :, c], my) for i in range(nx + 1)])
        assert rel_err(diffn[..., c], ref) < 1e-12


def _fields_xyz(nx, my, mz):
    i = np.arange(nx)[:, None, None]
    j = np.arange(my)[None, :, None]
    k = np.arange(mz)[None, None, :]
    return i, j, k


def test_compute_omega_with_x_variation(tmp_path, orc):
    """boundary_condition.F90:1648-1765: omega = curl u with central
    x-differences of ubcp over 2 dx (periodic, plane nx+1 a copy of plane 1)
    and exact spectral y/z derivatives.  Every component of u varies in x,
    y and z, so each of the six derivative terms is exercised."""
    st, o = _sphere(tmp_path, orc, "OM")
    nx, my, mz, dx = st.nx, st.my, st.mz, st.config.dx
    i, j, k = _fields_xyz(nx, my, mz)
    th, gy, gz = 2 * np.pi * 3 / nx, 2 * np.pi * 2 / my, 2 * np.pi / mz
    Ly, Lz = my * dx, mz * dx
    fy, fz = 2 * np.pi * 2 / Ly, 2 * np.pi / Lz  # physical wavenumbers of gy, gz
    # u = a cos(th i) cos(gy j) + b sin(gz k) cos(th i), etc.
    U = lambda ii: 0.3 * np.cos(th * ii) * np.cos(gy * j) + 0.2 * np.sin(gz * k) * np.cos(th * ii)
    V = lambda ii: 0.4 * np.sin(th * ii) + 0.1 * np.cos(th * ii) * np.sin(gz * k) + 0 * j
    W = lambda ii: 0.25 * np.cos(th * ii) * np.cos(gy * j) + 0.15 * np.sin(th * ii) + 0 * k
    dUdy = lambda ii: -0.3 * fy * np.cos(th * ii) * np.sin(gy * j) + 0 * k
    dUdz = lambda ii: 0.2 * fz * np.cos(gz * k) * np.cos(th * ii) + 0 * j
    dVdz = lambda ii: 0.1 * fz * np.cos(th * ii) * np.cos(gz * k) + 0 * j
    dWdy = lambda ii: -0.25 * fy * np.cos(th * ii) * np.sin(gy * j) + 0 * k
    ddx = lambda F: (F(i + 1) - F(i - 1)) / (2 * dx)
    u = np.zeros((nx + 1, my // 2 + 1, mz, 3), dtype=np.complex128)
    for c, F in enumerate((U, V, W)):
        u[:nx, :, :, c] = _spec_of(np.broadcast_to(F(i), (nx, my, mz)), nx)
    u[nx] = u[0]
    o.set_field("u", u)
    o.compute_new_timestep(1)
    o.nonlinear(1)
    om = o.field("omega")
    ref = np.zeros((nx + 1, my, mz, 3))
    ref[:nx, ..., 0] = dWdy(i) - dVdz(i)
    ref[:nx, ..., 1] = dUdz(i) - ddx(W)
    ref[:nx, ..., 2] = ddx(V) - dUdy(i)
    ref[nx] = ref[0]
    for c in range(3):
        assert rel_err(om[..., c], ref[..., c]) < 1e-12, c


def test_calc_pgrad_with_x_variation(tmp_path, orc):
    """boundary_condition.F90:1817-1896: the pressure gradient at the pressure
    points is the central x-difference of p over 2 dx (periodic over the nx
    pressure planes) and exact spectral y/z derivatives, plus mpg."""
    st, o = _sphere(tmp_path, orc, "PG")
    nx, my, mz, dx = st.nx, st.my, st.mz, st.config.dx
    i, j, k = _fields_xyz(nx, my, mz)
    th, gy, gz = 2 * np.pi * 2 / nx, 2 * np.pi * 3 / my, 2 * np.pi * 2 / mz
    fy, fz = gy / dx, gz / dx
    P = lambda ii: 0.5 * np.cos(th * ii + gy * j) + 0.3 * np.sin(gz * k) * np.cos(th * ii)
    dPdy = lambda ii: -0.5 * fy * np.sin(th * ii + gy * j) + 0 * k
    dPdz = lambda ii: 0.3 * fz * np.cos(gz * k) * np.cos(th * ii) + 0 * j
    o.set_field("p", _spec_of(np.broadcast_to(P(i), (nx, my, mz)), nx))
    s = o.scalars()
    s.umean[0] = 0.5  # away from the controller's target, so mpg is not zero
    o.set_scalars(s)
    o.compute_new_timestep(1)
    o.nonlinear(1)
    mpg = np.array(o.scalars().mpg)
    assert abs(mpg[0]) > 0
    ppr = o.field("ppr")
    ref = [(P(i + 1) - P(i - 1)) / (2 * dx), dPdy(i), dPdz(i)]
    for c in range(3):
        assert rel_err(ppr[..., c] - mpg[c], np.broadcast_to(ref[c], (nx, my, mz))) < 1e-12, c
        assert abs(np.mean(ppr[..., c]) - mpg[c]) < 1e-12 * max(1.0, abs(mpg[c]))


# ----------------------------------------------- reflection and mpg
def _uniform_flow(o, U):
    s = o.scalars()
    for n in range(3):
        s.umean[n] = U[n]
    o.set_scalars(s)


def test_reflected_velocity_of_a_sphere_at_rest(tmp_path, orc):
    """boundary_condition.F90:1329-1376.  In a uniform flow U past a sphere at
    rest (no pressure, no viscous or nonlinear terms, mpg held at zero) the
    reflection through the surface with symmetric dr (revernorm = 1) gives
    u_ref = -U at every outer point, so each reversal point carries
    F = cf (u_i - u_ref) dA2 dr dx = 2 cf U dA2 dr dx with
    dA2 = 4 pi (r_i dx)^2 / nrpr.  Summed and divided by dx^3 N (frmean),
    frmean = 2 cf U 4 pi (r_i dx)^2 dr dx / (dx^3 N)."""
    st, o = _sphere(tmp_path, orc, "RF", set_mpg=True, mpg=(0.0, 0.0, 0.0))
    U = (0.7, -0.2, 0.4)
    _uniform_flow(o, U)
    o.compute_new_timestep(1)
    o.nonlinear(1)
    s = o.scalars()
    dx, dr = st.config.dx, st.config.dr
    ri = st.radbdy[0] - dr
    N = st.nx * st.my * st.mz
    for n in range(3):
        ref = 2 * s.cf * U[n] * 4 * np.pi * (ri * dx) ** 2 * dr * dx / (dx ** 3 * N)
        assert abs(s.frmean[n] - ref) < 1e-12 * abs(ref), n
    assert np.max(np.abs(o.field("nlbcp"))) < 1e-14


@pytest.mark.parametrize("revernorm", [1, 0])
def test_sphere_moving_with_the_flow_is_not_forced(tmp_path, orc, revernorm):
    """The same flow with the sphere moving at U: the reflected velocity is U
    itself (tangential and, with revernorm = 1, normal part), so no point
    carries a force and fr stays zero."""
    st, o = _sphere(tmp_path, orc, f"MW{revernorm}", set_mpg=True, mpg=(0.0, 0.0, 0.0),
                    revernorm=revernorm)
    U = (0.7, -0.2, 0.4)
    _uniform_flow(o, U)
    o.update_particles(st.xc, np.array([U]))
    o.compute_new_timestep(1)
    o.nonlinear(1)
    s = o.scalars()
    scale = abs(s.cf) * 0.7
    fr = o.field("fr")
    if revernorm:
        assert np.max(np.abs(fr)) < 1e-12 * scale
        assert max(abs(x) for x in s.frmean) < 1e-14 * scale
    else:
        # only the tangential part is reflected, so u_i - u_ref is the normal
        # part n (n . U) of U at each point (n the unit normal of the point)
        assert np.max(np.abs(fr)) > 1e-3 * scale
        nrm = st.bndpts / np.sqrt((st.bndpts ** 2).sum(0))
        un = nrm @ (nrm.T @ np.array(U))
        dx, dr = st.config.dx, st.config.dr
        ri = st.radbdy[0] - dr
        da2 = 4 * np.pi * (ri * dx) ** 2 / st.nbnd
        N = st.nx * st.my * st.mz
        ref = s.cf * un * da2 * dr * dx / (dx ** 3 * N)
        assert rel_err(np.array(s.frmean), ref) < 1e-12


def test_mpg_controller_without_surface_drag(tmp_path, orc):
    """boundary_condition.F90:373-435 (fixed bed).  A uniform flow U with no
    pressure and no vorticity puts no drag on the surface, so the controller
    gives mpg = cf (ufmean_des - ufmean) with ufmean = U, and velstep's mean
    update umean += (-mpg + frmean) dt (steptheflow.F90:99) then moves the
    mean flow by dt cf (U - ufmean_des): all the way to ufmean_des when
    cf = -1/dt.  A surface force F on the bed (pressure drag of a linear
    pressure field, viscous drag of a shear mode) enters as -F / (voldom
    (1 - maxvolfrac)): the mean gradient that balances the drag over the
    fluid volume (pres_total accumulates +p n dA, minus the particles' pres,
    boundary_condition.F90:1613-1619)."""
    st, o = _sphere(tmp_path, orc, "MP")
    U = (0.3, 0.05, -0.02)
    _uniform_flow(o, U)
    o.compute_new_timestep(1)
    o.nonlinear(1)
    s = o.scalars()
    assert s.cf == -1.0 / s.dt
    f = o.forces()
    assert np.max(np.abs(f["pres"])) == 0.0 and np.max(np.abs(f["visc"])) < 1e-14
    for n in range(3):
        assert abs(s.ufmean[n] - U[n]) < 1e-12  # a sum over the box, rounded
        ref = s.cf * (s.ufmean_des[n] - s.ufmean[n])
        assert abs(s.mpg[n] - ref) < 1e-12 * max(abs(ref), 1e-3), n
    frm = np.array(s.frmean)
    o.velstep(1)
    s1 = o.scalars()
    for n in range(3):
        assert abs(s1.umean[n] - frm[n] * s.dt - (s.ufmean_des[n] + U[n] - s.ufmean[n])) < 1e-13, n

    # surface drag enters as minus the drag per unit fluid volume
    st2, o2 = _sphere(tmp_path, orc, "MQ")
    _uniform_flow(o2, U)
    o2.compute_new_timestep(1)
    G, dx = 0.7, st2.config.dx
    ii = np.arange(1, st2.nx + 1)
    o2.set_field("pbcp", np.broadcast_to((G * (ii + 0.5) * dx)[:, None, None], (st2.nx, st2.my, st2.mz)))
    o2.nonlinear(1)
    s2 = o2.scalars()
    pres = o2.forces()["pres"][0]
    assert pres[0] < 0
    vf = st2.config.voldom * (1.0 - s2.maxvolfrac)
    ref = s2.cf * (s2.ufmean_des[0] - s2.ufmean[0]) - pres[0] / vf
    assert abs(s2.mpg[0] - ref) < 1e-13 * abs(ref)
    assert abs(s2.mpg[0] - s.mpg[0]) > 1e-3 * abs(s.mpg[0])

    st3, o3 = _sphere(tmp_path, orc, "MV")
    my, mz, dx = st3.my, st3.mz, st3.config.dx
    A, k = 0.2, 2 * np.pi / (my * dx)
    plane = np.broadcast_to((A * np.cos(k * np.arange(my) * dx))[:, None], (my, mz))
    u = np.zeros((st3.nx + 1, my // 2 + 1, mz, 3), dtype=np.complex128)
    u[:, :, :, 0] = _r2c(plane)[None]
    o3.set_field("u", u)
    o3.compute_new_timestep(1)
    o3.nonlinear(1)
    s3 = o3.scalars()
    visc = o3.forces()["visc"][0]
    assert abs(visc[1]) > 0
    vf = st3.config.voldom * (1.0 - s3.maxvolfrac)
    for n in range(3):
        ref = s3.cf * (s3.ufmean_des[n] - s3.ufmean[n]) - visc[n] / vf
        assert abs(s3.mpg[n] - ref) < 1e-12 * max(abs(ref), abs(visc[1]) / vf), n


def test_surface_forcing_of_a_sphere_at_rest(tmp_path, orc):
    """calc_bnd_forcing (boundary_condition.F90:902-1102), DOBND = T with the
    reversal pairs off.  In a uniform flow U past a sphere at rest (no
    pressure, viscous or nonlinear terms, mpg held at zero) every surface
    point carries F = cf U dA1 dr dx with dA1 = 4 pi (r dx)^2 / nbnd, spread
    with its trilinear weights to all eight nodes of its stencil, fluid or
    solid.  So frmean = nbnd F/(dx^3 N) in closed form, and fr is the
    deposit of those point forces, formed here by brute force over the
    points."""
    st, o = _sphere(tmp_path, orc, "BF", set_mpg=True, mpg=(0.0, 0.0, 0.0), dorpr=False, dobnd=True)
    U = (0.7, -0.2, 0.4)
    _uniform_flow(o, U)
    o.compute_new_timestep(1)
    o.nonlinear(1)
    s = o.scalars()
    nx, my, mz, dx, dr = st.nx, st.my, st.mz, st.config.dx, st.config.dr
    rb = st.radbdy[0]
    N = nx * my * mz
    F = np.array([s.cf * U[n] * 4 * np.pi * (rb * dx) ** 2 / st.nbnd * dr * dx for n in range(3)])
    for n in range(3):
        ref = st.nbnd * F[n] / (dx ** 3 * N)
        assert abs(s.frmean[n] - ref) < 1e-12 * abs(ref), n
    dep = np.zeros((nx, my, mz, 3))
    pos = st.xc[0][:, None] + st.bndpts * rb
    for l in range(st.nbnd):
        p = pos[:, l]
        c = np.floor(p).astype(int)
        z = p - c
        for i in (0, 1):
            for j in (0, 1):
                for k in (0, 1):
                    w = (z[0] if i else 1 - z[0]) * (z[1] if j else 1 - z[1]) * (z[2] if k else 1 - z[2])
                    dep[(c[0] + i - 1) % nx, (c[1] + j - 1) % my, (c[2] + k - 1) % mz] += w * F / dx ** 3
    fr = o.field("fr")[:nx]  # fr - frmean
    assert rel_err(fr + np.array(s.frmean), dep) < 1e-12


def test_surface_forcing_of_a_sphere_moving_with_the_flow(tmp_path, orc):
    """The same flow with the sphere moving at U: u - linvel = 0 at every
    surface point, so calc_bnd_forcing deposits nothing."""
    st, o = _sphere(tmp_path, orc, "BM", set_mpg=True, mpg=(0.0, 0.0, 0.0), dorpr=False, dobnd=True)
    U = (0.7, -0.2, 0.4)
    _uniform_flow(o, U)
    o.update_particles(st.xc, np.array([U]))
    o.compute_new_timestep(1)
    o.nonlinear(1)
    s = o.scalars()
    assert np.max(np.abs(o.field("fr"))) < 1e-12 * abs(s.cf)
    assert max(abs(x) for x in s.frmean) < 1e-14 * abs(s.cf)


@pytest.mark.parametrize("rk3", [False, True])
def test_ferror_and_its_history(tmp_path, orc, rk3):
    """boundary_condition.F90:802-833.  Every bcset call sets
    ferror = |norm_drag - fold|/norm_drag and then fold = norm_drag, starting
    from fold = 0 (initialize_flo.F90:679-680), so the first ferror is exactly
    1; only the step's last stage enters its ferror at the head of
    ferror_array, whose older entries shift down by one (the oldest drops out),
    and ferror_hist is the mean of the array, summed from its head.  Checked
    stage by stage against the norm_drag sequence, with a history short
    enough (4) to wrap."""
    st, o = _sphere(tmp_path, orc, "FE", rk3=rk3)
    n = 4
    o.set_ferror_history(n)
    a, h = o.ferror_history(n)
    assert (a == 1.0).all() and h == 1.0
    assert o.scalars().fold == 0.0
    itrmax = 3 if rk3 else 1
    nd_prev, last = 0.0, []
    for step in range(6):
        for rks in range(1, itrmax + 1):
            o.compute_new_timestep(rks)
            o.nonlinear(rks)
            s = o.scalars()
            assert s.norm_drag > 0.0
            assert s.ferror == abs(s.norm_drag - nd_prev) / s.norm_drag
            assert s.fold == s.norm_drag
            nd_prev = s.norm_drag
            o.velstep(rks)
        last.append(s.ferror)
        want = (last[::-1] + [1.0] * n)[:n]
        a, h = o.ferror_history(n)
        assert list(a) == want, step
        acc = 0.0
        for v in want:
            acc += v
        assert h == acc / n
    with pytest.raises(solver.PibError):
        o.ferror_history(n + 1)


def test_ferror_of_a_bed_without_drag_is_one(tmp_path, orc):
    """boundary_condition.F90:807-811: norm_drag = 0 (a sphere moving with a
    uniform flow and no pressure carries no force) gives ferror = 1."""
    st, o = _sphere(tmp_path, orc, "FZ", set_mpg=True, mpg=(0.0, 0.0, 0.0))
    U = (0.7, -0.2, 0.4)
    _uniform_flow(o, U)
    o.update_particles(st.xc, np.array([U]))
    o.compute_new_timestep(1)
    o.nonlinear(1)
    s = o.scalars()
    assert s.norm_drag == 0.0
    assert s.ferror == 1.0 and s.fold == 0.0
