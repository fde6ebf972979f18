"""Closed-form pins of the oracle's whole scalar step (orc_scalstep: form_nlphi,
compute_phim_heat_ratio, calc_diffn, calc_bndrev_data, calc_inner_rev_data;
nl_allphi.F90:59-536, bc_scalar.F90:65-1033).  No GPU.

Each test fixes a piece by what the mathematics gives for a manufactured
input, so that a dropped term, a wrong sign, a heat-ratio factor on the
wrong side of the x wrap or a wrong plane fails it:
  - a y mode advected by a uniform v: nl = -i ky v phi exactly; diffn and
    grad phi of the mode are its spectral derivatives;
  - an x mode advected by a uniform u: the central x-flux gives the symbol
    sin(theta) sin(theta p)/dx, the upwind (MUSCL/Rusanov) flux the Fromm
    symbol, for u > 0 and u < 0;
  - phi = A exp(-lam x) with heat_ratio = exp(lam L) is what the twisted
    wrap phi(x + L) = phi(x)/heat_ratio holds exactly, so every wrap term
    (fdderiv(0), the slopes at the ends, calc_diffn's end planes) must give
    the derivative of the exponential; heat_ratio_new = heat_ratio/A;
  - a sphere whose surface value equals a uniform phi is not forced;
  - in a uniform phi0 the reversal forcing sums to cf (phi0 - phi_ref) 4 pi
    r_i^2 (the reflected value and the redistribution to solid nodes
    conserve it), also for a sphere across the x wrap (the two evaluations
    of a pair in cell nx); the DOBND surface forcing sums to cf (phi0 -
    phisurf) 4 pi r_b^2 less the shares the reference leaves on plane nx+1;
  - the surface flux of the exponential equals the trilinear interpolation
    of its exact x-gradient at the surface points, summed against the
    normals, computed here by brute force.
"""
import ctypes
import math

import numpy as np
import pytest

from conftest import make_setup, rel_err
from pureibm_ps2d_fd1d_amd import _abi, solver
from test_oracle import _bare_oracle, _get, _put

F_PHI, F_NL, F_ONL, F_FF = 14, 15, 16, 17


def _debug(lib, ctx, which, nx, my, mz):
    lib.orc_scalar_debug.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.POINTER(ctypes.c_double)]
    a = np.zeros((nx + 2, my, mz), order="F")
    assert lib.orc_scalar_debug(ctx, which, a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))) == 0
    return a


def _bare_flow(orc, nx, my, mz, dx, U, scheme=0, gamma=0.4, h=1.0):
    lib, ctx = _bare_oracle(orc, nx, my, mz, dx, 0.05, 0.01, scheme=scheme)
    s = _abi.PibScalars()
    assert lib.orc_get_scalars(ctx, ctypes.byref(s)) == 0
    for n in range(3):
        s.umean[n] = U[n]
    s.cf = -100.0
    assert lib.orc_set_scalars(ctx, ctypes.byref(s)) == 0
    assert lib.orc_compute_new_timestep(ctx, 1) == 0
    assert lib.orc_scalar_set(ctx, gamma, h, h) == 0
    assert lib.orc_scalar_surface(ctx, None) == 0
    return lib, ctx


def _ratios(lib, ctx):
    out = (ctypes.c_double * 4)()
    assert lib.orc_scalar_get(ctx, out) == 0
    return tuple(out)


@pytest.mark.parametrize("scheme", [_abi.PIB_SCHEME_CENTER, _abi.PIB_SCHEME_UPWIND])
def test_y_mode_in_a_uniform_flow(orc, scheme):
    nx, my, mz, dx, gamma = 12, 16, 8, 0.1, 0.4
    U, V, eps, jy = 0.8, -0.6, 0.3, 3
    lib, ctx = _bare_flow(orc, nx, my, mz, dx, (U, V, 0.0), scheme, gamma)
    shp = (nx + 1, my // 2 + 1, mz)
    phi = np.zeros(shp, dtype=np.complex128)
    phi[:, 0, 0] = 1.0
    phi[:, jy, 0] = eps / 2              # c2r doubles interior ky: phi = 1 + eps cos(ky y)
    _put(lib, ctx, F_PHI, phi)
    assert lib.orc_scalstep(ctx, 1) == 0
    nl = _get(lib, ctx, F_NL, shp)
    ky = 2 * math.pi * jy / (my * dx)
    ref = np.zeros(shp, dtype=np.complex128)
    ref[:, jy, 0] = -1j * ky * V * eps / 2
    assert np.max(np.abs(nl - ref)) < 1e-12 * abs(ky * V * eps)
    y = np.arange(my) * dx
    diffn = _debug(lib, ctx, 3, nx, my, mz)[1:nx + 2]
    gy = _debug(lib, ctx, 5, nx, my, mz)[1:nx + 2]
    gx, gz = _debug(lib, ctx, 4, nx, my, mz)[1:nx + 2], _debug(lib, ctx, 6, nx, my, mz)[1:nx + 2]
    cosy = np.broadcast_to(np.cos(ky * y)[None, :, None], diffn.shape)
    assert np.max(np.abs(diffn - gamma * (-ky ** 2) * eps * cosy)) < 1e-11 * gamma * ky ** 2 * eps
    assert np.max(np.abs(gy + ky * eps * np.sin(ky * y)[None, :, None])) < 1e-12 * ky * eps
    assert np.max(np.abs(gx)) < 1e-14 and np.max(np.abs(gz)) < 1e-12 * ky * eps
    hr, hrn, hro, g = _ratios(lib, ctx)
    assert abs(hr - 1.0) < 1e-14 and hro == 1.0 and g == gamma
    lib.orc_destroy(ctx)


def _fromm(theta, p, U, eps, dx):
    """-(D f)(p) of phi = 1 + eps cos(theta p) under the MUSCL/Rusanov x-flux
    with central slopes and a uniform u = U: the upwind side's linear
    reconstruction phi(p) +- (phi(p+1) - phi(p-1))/4 at each face."""
    e = np.exp(1j * theta * p)
    if U > 0:
        face = e * (1 + 0.5j * np.sin(theta))           # f(p+1/2)/(U eps) at the face right of p
    else:
        face = e * np.exp(1j * theta) * (1 - 0.5j * np.sin(theta))
    return -(U * eps / dx) * np.real(face * (1 - np.exp(-1j * theta)))


@pytest.mark.parametrize("scheme,U", [(_abi.PIB_SCHEME_CENTER, 0.7), (_abi.PIB_SCHEME_UPWIND, 0.7),
                                      (_abi.PIB_SCHEME_UPWIND, -0.5)])
def test_x_mode_in_a_uniform_flow(orc, scheme, U):
    nx, my, mz, dx, m, eps = 16, 8, 8, 0.1, 3, 0.25
    lib, ctx = _bare_flow(orc, nx, my, mz, dx, (U, 0.0, 0.0), scheme)
    theta = 2 * math.pi * m / nx
    p = np.arange(nx + 1)
    shp = (nx + 1, my // 2 + 1, mz)
    phi = np.zeros(shp, dtype=np.complex128)
    phi[:, 0, 0] = 1.0 + eps * np.cos(theta * p)
    _put(lib, ctx, F_PHI, phi)
    assert lib.orc_scalstep(ctx, 1) == 0
    nl = _get(lib, ctx, F_NL, shp)[:, 0, 0]
    if scheme == _abi.PIB_SCHEME_CENTER:
        ref = U * eps * np.sin(theta * p[:nx]) * np.sin(theta) / dx
    else:
        ref = _fromm(theta, p[:nx], U, eps, dx)
    assert np.max(np.abs(nl[:nx] - ref)) < 1e-11 * abs(U * eps / dx)
    assert abs(nl[nx] - nl[0]) < 1e-13 * abs(U * eps / dx)   # nlphif(mx) = nlphif(1)/heat_ratio, ratio 1
    lib.orc_destroy(ctx)


@pytest.mark.parametrize("scheme", [_abi.PIB_SCHEME_CENTER, _abi.PIB_SCHEME_UPWIND])
def test_twisted_exponential(orc, scheme):
    nx, my, mz, dx, gamma, U = 16, 8, 8, 0.1, 0.3, 0.9
    lam, A = 0.8, 1.3
    L = nx * dx
    h = math.exp(lam * L)
    lib, ctx = _bare_flow(orc, nx, my, mz, dx, (U, 0.0, 0.0), scheme, gamma, h)
    q = np.arange(-2, nx + 3)
    an = A * np.exp(-lam * q * dx)                   # phi of plane p = q (0-based), extended
    phi_p = lambda k: an[k + 2]
    shp = (nx + 1, my // 2 + 1, mz)
    phi = np.zeros(shp, dtype=np.complex128)
    phi[:, 0, 0] = phi_p(np.arange(nx + 1))
    _put(lib, ctx, F_PHI, phi)
    assert lib.orc_scalstep(ctx, 1) == 0
    nl = _get(lib, ctx, F_NL, shp)[:, 0, 0]
    p = np.arange(nx)
    if scheme == _abi.PIB_SCHEME_CENTER:
        ref = -U * (phi_p(p + 1) - phi_p(p - 1)) / (2 * dx)
    else:
        face = lambda k: U * (phi_p(k) + (phi_p(k + 1) - phi_p(k - 1)) / 4)   # U > 0: left state
        ref = -(face(p) - face(p - 1)) / dx
    scale = np.max(np.abs(ref))
    assert np.max(np.abs(nl[:nx] - ref)) < 1e-12 * scale
    assert abs(nl[nx] - nl[0] / h) < 1e-12 * scale
    # calc_diffn of planes 1..mx: diffn = gamma delta_x^2 phi, grad_x = central difference
    pp = np.arange(nx + 1)
    diffn = _debug(lib, ctx, 3, nx, my, mz)[1:nx + 2, 0, 0]
    gx = _debug(lib, ctx, 4, nx, my, mz)[1:nx + 2, 0, 0]
    d2 = gamma * (phi_p(pp + 1) - 2 * phi_p(pp) + phi_p(pp - 1)) / dx ** 2
    d1 = (phi_p(pp + 1) - phi_p(pp - 1)) / (2 * dx)
    assert rel_err(diffn, d2) < 1e-11 and rel_err(gx, d1) < 1e-12
    # compute_phim_heat_ratio: 1/phi of plane mx = h/A; rolled into heat_ratio
    hr, hrn, hro, _ = _ratios(lib, ctx)
    assert abs(hr - h / A) < 1e-13 * h and abs(hrn - h / A) < 1e-13 * h and hro == h
    lib.orc_destroy(ctx)


def test_heat_ratio_is_the_flow_weighted_mean(orc):
    """phim(mx) = sum u phi / sum u over the fluid of plane mx: with u = U + a
    cos(ky y) and phi = B + b cos(ky y), phim = B + a b/(2U)."""
    nx, my, mz, dx = 8, 16, 8, 0.1
    U, a, B, b, jy = 0.9, 0.3, 1.4, 0.5, 2
    lib, ctx = _bare_oracle(orc, nx, my, mz, dx, 0.05, 0.01)
    s = _abi.PibScalars()
    lib.orc_get_scalars(ctx, ctypes.byref(s))
    s.umean[0] = U
    lib.orc_set_scalars(ctx, ctypes.byref(s))
    shp3 = (nx + 1, my // 2 + 1, mz, 3)
    u = np.zeros(shp3, dtype=np.complex128)
    u[:, jy, 0, 0] = a / 2
    _put(lib, ctx, 0, u)
    assert lib.orc_compute_new_timestep(ctx, 1) == 0
    assert lib.orc_scalar_set(ctx, 0.2, 1.0, 1.0) == 0
    assert lib.orc_scalar_surface(ctx, None) == 0
    shp = (nx + 1, my // 2 + 1, mz)
    phi = np.zeros(shp, dtype=np.complex128)
    phi[:, 0, 0] = B
    phi[:, jy, 0] = b / 2
    _put(lib, ctx, F_PHI, phi)
    assert lib.orc_scalstep(ctx, 1) == 0
    hr = _ratios(lib, ctx)[0]
    assert abs(hr - 1.0 / (B + a * b / (2 * U))) < 1e-13
    lib.orc_destroy(ctx)


# ------------------------------------------------------- spheres
def _sphere(tmp_path, orc, name, xc=None, phi0=1.0, phisurf=0.3, h=1.0, **kw):
    st = make_setup(tmp_path, name, kind="simple", my=24, dbydx=6, re=20.0, **kw)
    o = solver.Solver(st, lib=orc.load(), prefix="orc_")
    if xc is not None:
        o.update_particles(np.asarray(xc, dtype=np.float64).reshape(1, 3))
    s = o.scalars()
    U = (0.6, 0.2, -0.1)
    for n in range(3):
        s.umean[n] = U[n]
    o.set_scalars(s)
    o.compute_new_timestep(1)
    o.scalar_set(0.3, h, h)
    o.scalar_surface(phisurf)
    nx, my, mz = st.nx, st.my, st.mz
    phi = np.zeros((nx + 1, my // 2 + 1, mz), dtype=np.complex128)
    phi[:, 0, 0] = phi0
    o.set_field("phi", phi)
    return st, o


def test_isothermal_sphere_is_not_forced(tmp_path, orc):
    """phi = phisurf = 1 everywhere (the value the reference starts from,
    scalar_init.F90:158-171, which also makes phim(mx) = 1 and so keeps
    heat_ratio at 1): no forcing, no flux, and the step keeps phi."""
    for dobnd in (False, True):
        st, o = _sphere(tmp_path / str(dobnd), orc, "ISO", phi0=1.0, phisurf=1.0, dobnd=dobnd)
        o.scalstep(1)
        ff = o.field("ffphi")
        fb, fg, pf, ps = o.scalar_flux()
        assert np.max(np.abs(ff)) < 1e-12 * abs(o.scalars().cf)
        assert np.max(np.abs(fb)) < 1e-12
        phi = o.field("phi")
        assert np.max(np.abs(phi[:, 0, 0] - 1.0)) < 1e-12 and np.max(np.abs(phi[:, 1:, :])) < 1e-12
        assert abs(pf - 1.0) < 1e-13 and abs(ps - 1.0) < 1e-13
        o.close()


def _ff_total(o, st):
    ff = o.field("ffphi")[: st.nx, 0, 0]
    return float(np.real(np.sum(ff))) * st.my * st.mz


@pytest.mark.parametrize("xc0", [None, 1.3])
def test_reversal_forcing_total(tmp_path, orc, xc0):
    """Uniform phi0, surface value phisurf: every active pair carries
    cf (phi0 - phi_ref) dA2 dx with phi_ref = (phisurf (ro - ri) - phi0
    (rb - ri))/(ro - rb) and dA2 = 4 pi (ri dx)^2 / nrpr_active, so the
    forcing integrates to cf (phi0 - phi_ref) 4 pi ri^2 (cell units)."""
    phi0, phisurf = 1.0, 0.3
    xc = None
    if xc0 is not None:
        st0 = make_setup(tmp_path / "g", "G", kind="simple", my=24, dbydx=6, re=20.0)
        xc = [xc0, st0.xc[0, 1], st0.xc[0, 2]]
    st, o = _sphere(tmp_path, orc, "RV", xc=xc, phi0=phi0, phisurf=phisurf)
    o.scalstep(1)
    rb = st.radbdy[0]
    dr = st.config.dr
    ri, ro = rb - dr, rb + dr
    phiref = (phisurf * (ro - ri) - phi0 * (rb - ri)) / (ro - rb)
    cf = o.scalars().cf
    ref = cf * (phi0 - phiref) * 4 * math.pi * ri ** 2
    assert abs(_ff_total(o, st) - ref) < 1e-11 * abs(ref)
    o.close()


@pytest.mark.parametrize("xc0", [None, 1.3])
def test_surface_forcing_total(tmp_path, orc, xc0):
    """DOBND with the pairs off: each surface point carries cf (phi0 -
    phisurf) dA1 dx spread to its eight nodes; the share of a node on plane
    nx+1 (a point in cell nx) stays on that plane, which is not transformed."""
    phi0, phisurf = 1.0, 0.3
    xc = None
    if xc0 is not None:
        st0 = make_setup(tmp_path / "g", "G", kind="simple", my=24, dbydx=6, re=20.0)
        xc = [xc0, st0.xc[0, 1], st0.xc[0, 2]]
    st, o = _sphere(tmp_path, orc, "BV", xc=xc, phi0=phi0, phisurf=phisurf, dorpr=False, dobnd=True)
    o.scalstep(1)
    nx = st.nx
    rb = st.radbdy[0]
    cf = o.scalars().cf
    xcv = np.asarray(xc if xc is not None else st.xc[0], dtype=np.float64)
    pos = xcv[:, None] + st.bndpts * rb
    cell = np.floor(pos[0]).astype(int)
    xl = pos[0].copy()
    xl[cell < 1] += nx
    cell[cell < 1] += nx
    xl[cell >= nx + 1] -= nx
    cell[cell >= nx + 1] -= nx
    kept = 1.0 - np.where(cell == nx, xl - cell, 0.0)
    ref = cf * (phi0 - phisurf) * 4 * math.pi * rb ** 2 * np.mean(kept)
    if xc0 is not None:
        assert np.mean(kept) < 0.999          # some shares are left on plane nx+1
    assert abs(_ff_total(o, st) - ref) < 1e-11 * abs(ref)
    o.close()


def test_surface_flux_of_a_twisted_exponential(tmp_path, orc):
    """flux_body = sum over the surface points of n . grad phi interpolated
    (bc_scalar.F90:514-550).  For phi = A exp(-lam x) under the twisted wrap,
    grad phi = (central difference in x, 0, 0) at every node, the wrap factors
    making the values those of the exponential at the nodes of the point's
    image in the box."""
    st0 = make_setup(tmp_path / "g", "G", kind="simple", my=24, dbydx=6, re=20.0)
    nx, dx = st0.nx, st0.config.dx
    lam, A = 0.6, 1.2
    h = math.exp(lam * nx * dx)
    xc = [1.3, st0.xc[0, 1], st0.xc[0, 2]]
    st, o = _sphere(tmp_path, orc, "FX", xc=xc, phi0=1.0, phisurf=0.3, h=h)
    ex = lambda i: A * np.exp(-lam * (i - 1) * dx)        # phi at x node i (reference index)
    phi = o.field("phi") * 0
    phi[:, 0, 0] = ex(np.arange(1, nx + 2))
    o.set_field("phi", phi)
    o.scalstep(1)
    fb = o.scalar_flux()[0]
    g = lambda i: (ex(i + 1) - ex(i - 1)) / (2 * dx)
    rb = st.radbdy[0]
    pos = np.asarray(xc)[:, None] + st.bndpts * rb
    tot = 0.0
    for l in range(st.nbnd):
        x = pos[0, l]
        c = math.floor(x)
        if c < 1:
            c, x = c + nx, x + nx
        if c >= nx + 1:
            c, x = c - nx, x - nx
        z = x - c
        tot += st.bndpts[0, l] * ((1 - z) * g(c) + z * g(c + 1))
    assert abs(fb[0] - tot) < 1e-11 * np.sum(np.abs(st.bndpts[0])) * abs(g(1))
    o.close()
