"""CPU checks that pin the oracle's scalar solve (phistep,
stepthescalar.F90:97-241) to what the mathematics fixes:

  - a cos/sin mode in x at heat_ratio = 1 is an eigenvector of the periodic
    second difference, so one Crank-Nicolson step multiplies it by
    (1 - dt c1 g lam)/(1 + dt c2 g lam), lam = 4 sin^2(pi m/nx)/dx^2 + w2;
  - an x-uniform mode with nlphif, onlphif, ffphi uniform gets the explicit
    terms dt c3 nl + dt c4 onl + dt ff with those exact weights;
  - phi_i = exp(kappa (i-1) dx), exp(kappa nx dx) = 1/h, satisfies the twisted
    wrap phi_(i+nx) = phi_i/h the solve imposes for heat_ratio = h, with the
    second-difference eigenvalue (2 cosh(kappa dx) - 2)/dx^2;
  - for random fields and heat_ratio != heat_ratio_new, the solution equals a
    dense solve of the cyclic system (corners a(1) h_new, c(nx)/h_new) built
    here independently of the Sherman-Morrison path;
  - the ratios roll (heat_ratio_old <- heat_ratio <- heat_ratio_new) as
    scalstep does after its mode loop (:89-90).
No GPU."""
import ctypes
import math

import numpy as np
import pytest

from pureibm_ps2d_fd1d_amd import _abi
from conftest import rel_err
from test_oracle import _bare_oracle, _get, _put

NX, MY, MZ, DX, DT = 16, 8, 8, 0.1, 0.01
F_PHI, F_NL, F_ONL, F_FF = 14, 15, 16, 17
SHP = (NX + 1, MY // 2 + 1, MZ)


def w2_grid():
    ky = 2 * math.pi * np.arange(MY // 2 + 1) / (MY * DX)
    kz = 2 * math.pi * np.fft.fftfreq(MZ, d=1.0 / MZ) / (MZ * DX)
    return ky[:, None] ** 2 + kz[None, :] ** 2


def scalar_oracle(orc, gamma, hr=1.0, hrn=None):
    lib, ctx = _bare_oracle(orc, NX, MY, MZ, DX, 0.05, DT)
    assert lib.orc_scalar_set(ctx, gamma, hr, hr if hrn is None else hrn) == 0
    return lib, ctx


def ratios(lib, ctx):
    out = (ctypes.c_double * 4)()
    assert lib.orc_scalar_get(ctx, out) == 0
    return tuple(out)


def test_cos_modes_get_the_crank_nicolson_factor(orc):
    rks = 1
    gamma = 0.7
    lib, ctx = scalar_oracle(orc, gamma)
    rng = np.random.default_rng(1)
    w2 = w2_grid()
    m = rng.integers(0, NX // 2 + 1, size=SHP[1:])
    amp = rng.standard_normal(SHP[1:]) + 1j * rng.standard_normal(SHP[1:])
    ph = rng.uniform(0, 2 * math.pi, size=SHP[1:])
    i = np.arange(NX + 1)[:, None, None]
    phi = amp * np.cos(2 * math.pi * m * i / NX + ph)
    _put(lib, ctx, F_PHI, phi)
    assert lib.orc_scalstep_modes(ctx, rks) == 0
    got = _get(lib, ctx, F_PHI, SHP)
    c1 = c2 = 0.5                                   # AB2/CN (initialize_flo.F90)
    lam = 4 * np.sin(math.pi * m / NX) ** 2 / DX ** 2 + w2
    g = (1 - DT * c1 * gamma * lam) / (1 + DT * c2 * gamma * lam)
    assert rel_err(got, g * phi) < 1e-12
    lib.orc_destroy(ctx)


def test_uniform_mode_explicit_weights(orc):
    gamma = 0.3
    lib, ctx = scalar_oracle(orc, gamma)
    w2 = w2_grid()
    phi0, nl, onl, ff = 0.8 - 0.1j, 1.3 + 0.4j, -0.6 + 2.0j, 0.25 - 0.5j
    for fid, v in ((F_PHI, phi0), (F_NL, nl), (F_ONL, onl), (F_FF, ff)):
        _put(lib, ctx, fid, np.full(SHP, v))
    assert lib.orc_scalstep_modes(ctx, 1) == 0
    got = _get(lib, ctx, F_PHI, SHP)
    c1, c2, c3, c4 = 0.5, 0.5, 1.5, -0.5
    want = (phi0 * (1 - DT * c1 * gamma * w2) + DT * c3 * nl + DT * c4 * onl + DT * ff) / (
        1 + DT * c2 * gamma * w2)
    assert rel_err(got, np.broadcast_to(want, SHP)) < 1e-13
    lib.orc_destroy(ctx)


@pytest.mark.parametrize("h", [0.8, 1.25])
def test_twisted_exponential_mode(orc, h):
    gamma = 0.5
    lib, ctx = scalar_oracle(orc, gamma, h, h)
    w2 = w2_grid()
    kappa = -math.log(h) / (NX * DX)
    i = np.arange(NX + 1)[:, None, None]
    phi = np.exp(kappa * i * DX) * (1.0 + 0.5j) * np.ones(SHP)
    _put(lib, ctx, F_PHI, phi)
    assert lib.orc_scalstep_modes(ctx, 1) == 0
    got = _get(lib, ctx, F_PHI, SHP)
    mu = (2 * math.cosh(kappa * DX) - 2) / DX ** 2
    c1 = c2 = 0.5
    g = (1 - DT * c1 * gamma * (w2 - mu)) / (1 + DT * c2 * gamma * (w2 - mu))
    assert rel_err(got, g * phi) < 1e-12
    assert rel_err(got[NX], got[0] / h) < 1e-15     # phif(mx) = phif(1)/heat_ratio_new
    hr, hrn, hro, gm = ratios(lib, ctx)
    assert (hr, hrn, hro, gm) == (h, h, h, gamma)
    lib.orc_destroy(ctx)


def test_random_fields_match_a_dense_cyclic_solve(orc):
    gamma, hr, hrn = 0.9, 1.1, 0.85
    lib, ctx = scalar_oracle(orc, gamma, hr, hrn)
    rng = np.random.default_rng(7)
    f = {fid: rng.standard_normal(SHP) + 1j * rng.standard_normal(SHP)
         for fid in (F_PHI, F_NL, F_ONL, F_FF)}
    for fid, v in f.items():
        _put(lib, ctx, fid, v)
    assert lib.orc_scalstep_modes(ctx, 1) == 0
    got = _get(lib, ctx, F_PHI, SHP)
    c1, c2, c3, c4 = 0.5, 0.5, 1.5, -0.5
    w2 = w2_grid()
    p = f[F_PHI]
    for j in range(SHP[1]):
        for k in range(SHP[2]):
            e = DT * gamma * c2 / DX ** 2
            A = np.diag(np.full(NX, 1 + DT * gamma * c2 * (2 / DX ** 2 + w2[j, k])))
            A += np.diag(np.full(NX - 1, -e), 1) + np.diag(np.full(NX - 1, -e), -1)
            A[0, NX - 1] = -e * hrn                    # phi_0 = h_new phi_nx
            A[NX - 1, 0] = -e / hrn                    # phi_(nx+1) = phi_1/h_new
            q = p[:, j, k]
            lap = np.empty(NX, dtype=complex)
            lap[1:] = q[:NX - 1] + q[2:NX + 1] - 2 * q[1:NX]
            lap[0] = q[1] - 2 * q[0] + hr * q[NX - 1]
            rhs = (q[:NX] * (1 - DT * c1 * gamma * w2[j, k]) + DT * c3 * f[F_NL][:NX, j, k]
                   + DT * f[F_FF][:NX, j, k] + DT * c4 * f[F_ONL][:NX, j, k]
                   + DT * gamma * c1 * lap / DX ** 2)
            want = np.linalg.solve(A, rhs)
            assert rel_err(got[:NX, j, k], want) < 1e-12, (j, k)
            assert abs(got[NX, j, k] - got[0, j, k] / hrn) <= 1e-15 * abs(got[0, j, k] / hrn)
    assert ratios(lib, ctx)[:3] == (hrn, hrn, hr)
    lib.orc_destroy(ctx)


def test_scalar_errors(orc):
    lib, ctx = _bare_oracle(orc, NX, MY, MZ, DX, 0.05, DT)
    assert lib.orc_scalstep_modes(ctx, 1) == _abi.PIB_ERR_STATE
    a = np.zeros(SHP, dtype=np.complex128, order="F")
    ptr = a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    assert lib.orc_get_field(ctx, F_PHI, ptr) == _abi.PIB_ERR_STATE
    assert lib.orc_set_field(ctx, F_NL, ptr) == _abi.PIB_ERR_STATE
    assert lib.orc_scalar_get(ctx, (ctypes.c_double * 4)()) == _abi.PIB_ERR_STATE
    assert lib.orc_get_field(ctx, 99, ptr) == _abi.PIB_ERR_ARG
    assert lib.orc_scalar_set(ctx, 0.0, 1.0, 1.0) == _abi.PIB_ERR_ARG
    assert lib.orc_scalar_set(ctx, 1.0, 1.0, 1.0) == 0
    assert lib.orc_scalstep_modes(ctx, 2) == _abi.PIB_ERR_ARG    # AB2 has one stage
    lib.orc_destroy(ctx)
