"""The adjoint of the fixed-bed step in the oracle (CPU), pinned by what the
mathematics fixes, not by a restatement of it:

* the step is a quadratic polynomial of its state (u, p, nl, umean) for a
  fixed bed and central x-fluxes, and the drag is linear in it, so a central
  difference of  <S(x), y> + wforce . force(x)  along any direction is exact
  up to rounding and must equal  <direction, S^T y>  from one adjoint step;
* dJ/dxc and dJ/dmpg of the drag after several checkpointed steps must equal
  central finite differences of the forward runs (h = 1e-6 cell), with no
  mask, reversal-pair or stencil-cell change within +-h;
* the gradient does not depend on the checkpoint spacing.
A planted mistake in the oracle's transposes (a sign, a dropped term, a
transposed operand) fails the first check.
"""
import ctypes
import os

import numpy as np
import pytest

from pureibm_ps2d_fd1d_amd import _abi, adjoint, configs, solver


def _setup(tmp_path, name="ADJ", my=16, dbydx=6, nbody=2, seed=3, **kw):
    d = str(tmp_path)
    configs.random_bed(d, name, my=my, dbydx=dbydx, nbody=nbody, seed=seed, re=20.0, adaptive=False,
                       set_mpg=True, mpg=(0.3, 0.05, -0.02), **kw)
    return solver.RunSetup(os.path.join(d, name))


def _dot(a, b):
    return float(np.sum((np.conj(a) * b).real))


def transpose_check(s, rng, nx, nsteps_before=3):
    """<dS, y> by a central difference against <dx, S^T y> from adjoint_step."""
    sc = s.scalars()
    for n in range(3):
        sc.umean[n] = sc.ufmean_des[n]
    s.set_scalars(sc)
    s.step(nsteps_before)
    u0, p0, nl0 = s.field("u"), s.field("p"), s.field("nl")
    s0 = s.scalars()

    def rc(shape):
        return rng.standard_normal(shape) + 1j * rng.standard_normal(shape)

    def periodic(a):
        a = a.copy()
        a[nx] = a[0]
        return a

    du, dp, dnl = periodic(rc(u0.shape)) * 1e-2, rc(p0.shape) * 1e-2, periodic(rc(nl0.shape)) * 1e-2
    dum = rng.standard_normal(3) * 1e-2
    yu, yp, ynl = rc(u0.shape), rc(p0.shape), rc(nl0.shape)
    yum = rng.standard_normal(3)
    wf = np.asfortranarray(rng.standard_normal((s.setup.nbody, 3)))

    def restore(eps):
        s.set_field("u", u0 + eps * du)
        s.set_field("p", p0 + eps * dp)
        s.set_field("nl", nl0 + eps * dnl)
        sc = s.scalars()
        ctypes.memmove(ctypes.byref(sc), ctypes.byref(s0), ctypes.sizeof(sc))
        for n in range(3):
            sc.umean[n] = s0.umean[n] + eps * dum[n]
        s.set_scalars(sc)

    def F(eps):
        restore(eps)
        s.step(1)
        u, p, nl = s.field("u"), s.field("p"), s.field("nl")
        um = np.array(s.scalars().umean)
        return (_dot(u[:nx], yu[:nx]) + _dot(p, yp) + _dot(nl[:nx], ynl[:nx]) + um @ yum
                + float(np.sum(wf * s.forces()["force"])))

    eps = 1e-3
    fd = (F(eps) - F(-eps)) / (2 * eps)
    restore(0.0)
    s.adjoint_begin()
    s.set_field("au", yu)
    s.set_field("ap", yp)
    s.set_field("anl", ynl)
    s.adjoint_set_umean(yum)
    s.adjoint_step(wf)
    au, ap, anl = s.field("au"), s.field("ap"), s.field("anl")
    _, _, aum = s.adjoint_get()
    ad = _dot(du[:nx], au[:nx]) + _dot(dp, ap) + _dot(dnl[:nx], anl[:nx]) + dum @ aum
    return fd, ad


def test_adjoint_step_is_the_transpose_of_the_step(tmp_path, orc):
    st = _setup(tmp_path)
    s = solver.Solver(st, lib=orc.load(), prefix="orc_")
    fd, ad = transpose_check(s, np.random.default_rng(5), s.nx)
    print(f"central difference {fd:.15e}, adjoint {ad:.15e}")
    assert abs(fd - ad) <= 1e-10 * abs(fd)


def test_adjoint_step_transpose_revernorm0_and_ramp(tmp_path, orc):
    st = _setup(tmp_path, revernorm=0, ramp_frac_steps=2, dr=0.8)
    s = solver.Solver(st, lib=orc.load(), prefix="orc_")
    fd, ad = transpose_check(s, np.random.default_rng(7), s.nx, nsteps_before=1)
    assert abs(fd - ad) <= 1e-10 * abs(fd)


def point_cells_move(setup, xc, m, d, h):
    """True if any surface, inner, outer or pressure point of body m changes
    its stencil cell when xc[m, d] moves by +-h."""
    dr = setup.config.dr
    rb = setup.radbdy[m]
    pos = []
    for r in (rb, rb - dr, rb + dr):
        p = xc[m][:, None] + setup.bndpts * r
        pos.append(p)
        q = p.copy()
        q[0] -= 0.5
        pos.append(q)
    c = np.concatenate([p[d] for p in pos])
    return bool(np.any(np.floor(c + h) != np.floor(c)) or np.any(np.floor(c - h) != np.floor(c)))


def fd_gradient(P, setup, xc0, mpg0, h, coords=None):
    """Central differences of P.objective in xc (coords: list of (m, d)) and
    mpg; coordinates whose points change stencil cells are returned in skip."""
    nb = setup.nbody
    coords = [(m, d) for m in range(nb) for d in range(3)] if coords is None else coords
    fd = np.full((nb, 3), np.nan)
    skip = []
    for (m, d) in coords:
        if point_cells_move(setup, xc0, m, d, h):
            skip.append((m, d))
            continue
        xp = xc0.copy()
        xp[m, d] += h
        Jp = P.objective(xp, mpg0, keep=False)
        cp, nrp = P.s.count_solid, P.s.nrpr_active.copy()
        xm = xc0.copy()
        xm[m, d] -= h
        Jm = P.objective(xm, mpg0, keep=False)
        cm, nrm = P.s.count_solid, P.s.nrpr_active.copy()
        if cp != cm or not np.array_equal(nrp, nrm):
            skip.append((m, d))
            continue
        fd[m, d] = (Jp - Jm) / (2 * h)
    fdm = np.zeros(3)
    for d in range(3):
        mp, mm = mpg0.copy(), mpg0.copy()
        mp[d] += h
        mm[d] -= h
        fdm[d] = (P.objective(xc0, mp, keep=False) - P.objective(xc0, mm, keep=False)) / (2 * h)
    return fd, fdm, skip


def test_gradient_matches_finite_differences(tmp_path, orc):
    st = _setup(tmp_path)
    P = adjoint.DragProblem(st, nsteps=5, every=2, lib=orc.load(), prefix="orc_")
    xc0, mpg0 = st.xc.copy(), np.array([0.3, 0.05, -0.02])
    P.objective(xc0, mpg0)
    dxc, dmpg = P.gradient()
    fd, fdm, skip = fd_gradient(P, st, xc0, mpg0, 1e-6)
    ok = ~np.isnan(fd)
    print("adjoint", dxc[ok], "\nfd", fd[ok], "\nskipped", skip)
    assert ok.sum() >= 5
    assert np.max(np.abs(dxc[ok] - fd[ok])) <= 1e-6 * np.max(np.abs(fd[ok]))
    assert np.max(np.abs(dmpg - fdm)) <= 1e-6 * np.max(np.abs(fdm))
    P.close()


def test_gradient_does_not_depend_on_the_checkpoint_spacing(tmp_path, orc):
    st = _setup(tmp_path)
    g = []
    for every in (1, 3, 7):
        P = adjoint.DragProblem(st, nsteps=7, every=every, lib=orc.load(), prefix="orc_")
        P.objective()
        g.append(P.gradient())
        P.close()
    for dxc, dmpg in g[1:]:
        assert np.array_equal(dxc, g[0][0]) and np.array_equal(dmpg, g[0][1])


def test_torch_function_returns_the_adjoint_gradient(tmp_path, orc):
    torch = pytest.importorskip("torch")
    st = _setup(tmp_path)
    P = adjoint.DragProblem(st, nsteps=4, every=2, lib=orc.load(), prefix="orc_")
    xc = torch.tensor(st.xc, dtype=torch.float64, requires_grad=True)
    mpg = torch.tensor([0.3, 0.05, -0.02], dtype=torch.float64, requires_grad=True)
    J = adjoint.drag_objective().apply(xc, mpg, P)
    J.backward()
    P2 = adjoint.DragProblem(st, nsteps=4, every=2, lib=orc.load(), prefix="orc_")
    J2 = P2.objective(st.xc, np.array([0.3, 0.05, -0.02]))
    dxc, dmpg = P2.gradient()
    assert float(J.detach()) == J2
    assert np.array_equal(xc.grad.numpy(), dxc) and np.array_equal(mpg.grad.numpy(), dmpg)
    P.close()
    P2.close()


def test_unsupported_runs_are_refused(tmp_path, orc):
    d = str(tmp_path)
    configs.random_bed(d, "ADA", my=16, dbydx=6, nbody=2, seed=3, re=20.0, adaptive=True, set_mpg=True)
    s = solver.Solver(solver.RunSetup(os.path.join(d, "ADA")), lib=orc.load(), prefix="orc_")
    s.adjoint_begin()
    with pytest.raises(solver.PibError) as e:
        s.adjoint_step(None)
    assert e.value.code == _abi.PIB_ERR_UNSUPPORTED
    with pytest.raises(solver.PibError) as e:
        s.checkpoint_pop()
    assert e.value.code == _abi.PIB_ERR_STATE
