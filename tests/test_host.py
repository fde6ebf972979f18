"""Host side (Fortran library): the reference's namelist input contract and
the initflo-derived run constants; the C-ABI export table."""
import ctypes
import math
import os
import re
import shutil

import numpy as np
import pytest

from conftest import ROOT, make_setup
from pureibm_ps2d_fd1d_amd import _abi, configs, solver

GOLD = os.path.join(ROOT, "tests", "golden")


@pytest.mark.parametrize("deck", ["refsample", "refmis1"])
def test_reads_reference_example_decks(tmp_path, deck):
    """The reference's own example inputs parse (all groups, with the
    reference's variables); what they ask for beyond the host (the 'random'
    configuration generator; IRESTART = 1 at configuration, where a restart
    is read by pib_host_read_restart instead) is refused cleanly.  The MIS
    deck's ISCALON = 1 with IMOVE = 1 passes that check."""
    shutil.copy(os.path.join(GOLD, f"{deck}_floparam.in"), tmp_path / f"{deck}_floparam.in")
    host = _abi.load_host()
    p = str(tmp_path / f"{deck}_floparam.in").encode()
    assert host.pib_host_read(p, len(p)) == 0
    rn = str(tmp_path / deck).encode()
    st = host.pib_host_configure(rn, len(rn))
    assert st == _abi.PIB_ERR_UNSUPPORTED
    assert "IRESTART" in _abi.host_error() or "random" in _abi.host_error()
    assert "ISCALON" not in _abi.host_error()


def test_missing_deck_is_an_error(tmp_path):
    host = _abi.load_host()
    p = str(tmp_path / "nope_floparam.in").encode()
    assert host.pib_host_read(p, len(p)) == _abi.PIB_ERR_ARG


def test_simple_configuration_constants(tmp_path):
    """'simple' (initialize_flo.F90:1372-1424): one sphere at (mx1/2+1,
    my/2+1, mz/2+1), radbdy = my/(2 lybyd), dx = lybyd*dia/my."""
    st = make_setup(tmp_path, "S", kind="simple", my=32, dbydx=8, re=20.0)
    assert st.nbody == 1 and (st.nx, st.my, st.mz) == (32, 32, 32)
    assert np.allclose(st.xc, [[17.0, 17.0, 17.0]])
    assert abs(st.radbdy[0] - 4.0) < 1e-12
    c = st.config
    assert abs(c.dx - 4.0 / 32) < 1e-15
    assert abs(c.ucharmod - 20.0 * 0.012 / 1.0) < 1e-15
    assert abs(c.voldom - 4.0 ** 3) < 1e-12
    assert st.nbnd == 1158 * 0 + st.nbnd  # point count depends on r = 4


def test_default_configuration_reads_centres(tmp_path):
    centers = np.array([[0.5, 1.25, 2.0], [2.5, 2.5, 0.75]])
    configs.write_case(str(tmp_path), "D", my=24, lybyd=3.0, dbydx=8.0, centers=centers)
    st = solver.RunSetup(str(tmp_path / "D"))
    assert st.nbody == 2
    assert np.allclose(st.xc, centers * 8.0)
    assert np.allclose(st.radbdy, 4.0)


def test_initial_time_step_and_mean_flow(tmp_path):
    """initflo (initialize_flo.F90:333-359, 552-672): ufmean_des = uchar/(1-phi),
    umean = (1-phi) ufmean_des, dt = cfl*dx/(|uchar|/(1-phi)), cf = -1/dt."""
    st = make_setup(tmp_path, "T", kind="simple", my=32, dbydx=8, re=20.0, cfl=0.2)
    count_solid = 251
    s = st.finish(count_solid)
    phi = count_solid / 32 ** 3
    uchar = 20.0 * 0.012
    assert abs(s.maxvolfrac - phi) < 1e-15
    assert abs(s.ufmean_des[0] - uchar / (1 - phi)) < 1e-15
    assert abs(s.umean[0] - (1 - phi) * uchar / (1 - phi)) < 1e-15
    dt = 0.2 * (4.0 / 32) / (uchar / (1 - phi))
    assert abs(s.dt - dt) < 1e-15 * dt
    assert abs(s.cf + 1.0 / dt) < 1e-12 / dt
    assert s.ramp_frac_time == 0.0


def test_unsupported_options_are_refused(tmp_path):
    configs.write_case(str(tmp_path), "U", my=16, lybyd=2.0, dbydx=8.0, input_type="simple",
                       vol_frac1=0.06, discr_scheme="quick")
    with pytest.raises(solver.PibError) as e:
        solver.RunSetup(str(tmp_path / "U"))
    assert e.value.code == _abi.PIB_ERR_UNSUPPORTED


def test_gpu_library_exports_the_header():
    """Every entry point declared in include/*.h is exported by the HIP
    library (loading it makes no GPU call)."""
    names = set()
    for h in ("pureibm_gpu.h", "pureibm_dem.h"):
        hdr = open(os.path.join(ROOT, "include", h)).read()
        names |= set(re.findall(r"\b(pib_[a-z0-9_]+)\s*\(", hdr))
    assert len(names) >= 26
    lib = ctypes.CDLL(_abi.gpu_lib_path())
    for n in sorted(names):
        assert hasattr(lib, n), n
    assert set(_abi.GPU_SYMBOLS) <= names


def test_fortran_mirrors_match_the_c_structs():
    """The bind(C) types of fsrc/pib_types.f90 have the C structs' sizes."""
    sizes = (ctypes.c_int * 4)()
    assert _abi.load_host().pib_host_struct_sizes(sizes) == 0
    want = [ctypes.sizeof(_abi.PibConfig), ctypes.sizeof(_abi.PibScalars),
            ctypes.sizeof(_abi.PibDemConfig), ctypes.sizeof(_abi.PibDemInfo)]
    assert list(sizes) == want


FORTRAN_CALLER = """
program caller
  use pib_types
  use iso_c_binding
  implicit none
  type(c_ptr) :: ctx
  complex(c_double_complex) :: u(9, 5, 8, 3), phif(9, 5, 8, 1), bc(5, 8)
  real(c_double) :: ubcp(9, 8, 8, 3), ar(8, 8), aa(16), x(16), out4(4)
  integer(c_int) :: st
  ctx = c_null_ptr
  st = pib_set_field(ctx, PIB_F_U, u)
  st = pib_get_field(ctx, 5, ubcp)
  st = pib_scalar_set(ctx, 0.1d0, 1.0d0, 1.0d0)
  st = pib_set_field(ctx, PIB_F_PHI, phif(:, :, :, 1))
  st = pib_scalstep_modes(ctx, 1)
  st = pib_get_field(ctx, PIB_F_PHI, phif(:, :, :, 1))
  st = pib_scalar_get(ctx, out4)
  st = pib_ff2rc_planes(ctx, ar, 64_c_long_long, 1, bc, 40_c_long_long)
  st = pib_ff2cr_planes(ctx, bc, 40_c_long_long, 1, ar, 64_c_long_long)
  st = pib_tridiag_batched(ctx, 16, 1, aa, aa, aa, aa, x, 1)
end program caller
"""

# the adjoint sweep of the Fortran binding example: forward run checkpointed
# every 8 steps, then each segment recorded and reversed
FORTRAN_ADJOINT = """
program adjoint_sweep
  use iso_c_binding
  use pib_types
  implicit none
  integer, parameter :: nbody = 2, nseg = 4, every = 8
  type(c_ptr) :: ctx
  integer(c_int) :: st, k, j
  real(c_double), target :: wforce(nbody, 3)
  real(c_double) :: dxc(nbody, 3), dmpg(3), dumean(3)
  ctx = c_null_ptr
  do k = 1, nseg
     st = pib_checkpoint_push(ctx)
     st = pib_checkpoint_steps(ctx, every)
  end do
  st = pib_adjoint_begin(ctx)
  dumean = 0
  st = pib_adjoint_set_umean(ctx, dumean)
  do k = nseg, 1, -1
     st = pib_checkpoint_pop(ctx)
     if (k == nseg) then          ! the drag of the last step is in J
        st = pib_adjoint_record(ctx, every - 1)
        st = pib_adjoint_step(ctx, c_loc(wforce))
        do j = every - 1, 1, -1
           st = pib_adjoint_step(ctx, c_null_ptr)
        end do
     else
        st = pib_adjoint_record(ctx, every)
        do j = every, 1, -1
           st = pib_adjoint_step(ctx, c_null_ptr)
        end do
     end if
  end do
  st = pib_adjoint_get(ctx, dxc, dmpg, dumean)
end program adjoint_sweep
"""

# the fixed-bed and moving-bed call sequences of the Fortran binding example
FORTRAN_STEP = """
program moving_step
  use iso_c_binding
  use pib_types
  implicit none
  integer, parameter :: nbody = 2, nbnd = 10, mx = 9, my = 8, mz = 8, my2 = 5
  type(c_ptr)        :: ctx, dem
  type(pib_config)   :: cfg
  type(pib_scalars)  :: sc
  type(pib_dem_config) :: dc
  type(pib_dem_info)   :: di
  integer(c_int)     :: st, rks, itrmax
  integer(c_long_long) :: nsolid
  integer(c_int), allocatable :: nrpr_active(:)
  real(c_double) :: bndpts(3, nbnd), cd(nbnd), radbdy(nbody), dt
  real(c_double), target :: xc(nbody, 3), velbdy(nbody, 3), angv(nbody, 3)
  real(c_double) :: force(nbody, 3), pres(nbody, 3), visc(nbody, 3), torq(nbody, 3)
  complex(c_double_complex) :: u(mx, my2, mz, 3)
  ctx = c_null_ptr
  cfg%nx = mx - 1;  cfg%my = my;  cfg%mz = mz;  cfg%dx = 0.1d0;  cfg%vis = 0.01d0
  cfg%precision = PIB_FP64
  st = pib_create(cfg, ctx)
  st = pib_set_points(ctx, nbnd, bndpts, cd)
  allocate(nrpr_active(nbody))
  st = pib_set_particles(ctx, nbody, xc, radbdy, c_loc(velbdy), c_loc(angv), nsolid, nrpr_active)
  st = pib_set_field(ctx, PIB_F_U, u)
  st = pib_set_scalars(ctx, sc)
  rks = 1
  st = pib_compute_new_timestep(ctx, rks)
  st = pib_nonlinear(ctx, rks)
  st = pib_velstep(ctx, rks)
  if (st == PIB_ERR_BLOWUP) stop
  st = pib_get_scalars(ctx, sc)
  st = pib_step(ctx, 3)
  dc%nbody = nbody;  dc%kn = 0
  st = pib_dem_create(dc, xc, radbdy, c_loc(velbdy), dem)
  st = pib_get_forces(ctx, force, pres, visc, torq)
  st = pib_dem_march(dem, dt, pres, visc, sc%mpg, sc%frame_accln, xc, velbdy, angv)
  if (st == PIB_ERR_MOTION) stop 'MOVEMENT UNDESIRED'
  itrmax = 1
  do rks = 1, itrmax
     st = pib_compute_new_timestep(ctx, rks)
     st = pib_nonlinear_begin(ctx, rks)
     st = pib_dem_march(dem, sc%dt, pres, visc, sc%mpg, sc%frame_accln, xc, velbdy, angv)
     st = pib_set_particles(ctx, nbody, xc, radbdy, c_loc(velbdy), c_loc(angv), nsolid, nrpr_active)
     st = pib_nonlinear_finish(ctx, rks)
     st = pib_velstep(ctx, rks)
  end do
  st = pib_dem_get_state(dem, xc, velbdy, angv, force, di)
end program moving_step
"""


def test_fortran_bindings_take_the_reference_arrays(tmp_path):
    """A Fortran caller passes the reference's complex(8) and real(8) arrays
    to the field, scalar, transform and tridiagonal entry points, and runs the
    fixed-bed, moving-bed and adjoint call sequences, through the interfaces
    of fsrc/pib_types.f90 (compile only, no call)."""
    fc = shutil.which("amdflang") or "/opt/rocm/bin/amdflang"
    mods = os.path.join(ROOT, "pureibm_ps2d_fd1d_amd", "_build")
    if not os.path.exists(fc) or not os.path.exists(os.path.join(mods, "pib_types.mod")):
        pytest.skip("amdflang or the built pib_types module is missing")
    import subprocess
    for name, text in (("caller", FORTRAN_CALLER), ("moving_step", FORTRAN_STEP), ("adjoint", FORTRAN_ADJOINT)):
        src = tmp_path / (name + ".f90")
        src.write_text(text)
        r = subprocess.run([fc, "-c", "-J", str(tmp_path), "-I", mods, str(src), "-o", str(tmp_path / (name + ".o"))],
                           capture_output=True, text=True)
        assert r.returncode == 0, (name, r.stderr)


C_CALLER = r"""
#include "pureibm_gpu.h"
#include "pureibm_dem.h"
#include <stddef.h>
int main(void) {
  void* ctx = NULL;
  void* dem = NULL;
  pib_config c = {0};
  pib_dem_config dc = {0};
  double out[4];
  int st = pib_create(&c, &ctx);
  st |= pib_scalar_set(ctx, 0.1, 1.0, 1.0);
  st |= pib_scalar_get(ctx, out);
  st |= pib_scalstep_modes(ctx, 1);
  st |= pib_set_ferror_history(ctx, PIB_NERR_DEFAULT, NULL);
  st |= pib_get_ferror_history(ctx, PIB_NERR_DEFAULT, NULL, out);
  st |= pib_dem_create(&dc, NULL, NULL, NULL, &dem);
  return st;
}
"""


def test_c_headers_compile_as_c99_and_link(tmp_path):
    """include/*.h are plain C99 (no warnings under -pedantic) and a C caller
    links against the HIP library (built, not run)."""
    cc = shutil.which("gcc")
    if not cc or not os.path.exists(_abi.gpu_lib_path()):
        pytest.skip("gcc or the built library is missing")
    src = tmp_path / "caller.c"
    src.write_text(C_CALLER)
    import subprocess
    r = subprocess.run([cc, "-std=c99", "-Wall", "-Wextra", "-pedantic", "-Werror", "-I",
                        os.path.join(ROOT, "include"), str(src), "-L", os.path.dirname(_abi.gpu_lib_path()),
                        "-lpureibm_gpu", "-o", str(tmp_path / "caller")], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr


def test_product_never_imports_the_oracle():
    pkg = os.path.join(ROOT, "pureibm_ps2d_fd1d_amd")
    for dirpath, _, files in os.walk(pkg):
        for f in files:
            if f.endswith((".py", ".hip", ".h", ".f90", ".cpp")):
                txt = open(os.path.join(dirpath, f)).read()
                assert "import oracle" not in txt and "liboracle" not in txt, f


def test_batched_realizations_are_reproducible(tmp_path):
    """The batched workload's decks are fixed by their seeds: grid sizes cycle
    96/128/160 and the sphere sets repeat exactly."""
    from pureibm_ps2d_fd1d_amd import configs
    sizes = []
    for k in range(6):
        b1 = configs.c3_realization(str(tmp_path / "a"), k)
        b2 = configs.c3_realization(str(tmp_path / "b"), k)
        t1 = open(b1 + "_sphr_center.inp").read()
        t2 = open(b2 + "_sphr_center.inp").read()
        assert t1 == t2
        nl = open(b1 + "_floparam.in").read()
        sizes.append(int(nl.split("MY = ")[1].split(",")[0]))
    assert sizes == [96, 128, 160, 96, 128, 160]


def test_scalar_deck_is_configured(tmp_path):
    """An ISCALON = 1 deck: gamma = vis/Pr (scalar_init.F90:73), the initial
    dt takes the diffusive limit dx^2 (1 - maxvolfrac)/gamma beside the
    convective dx/umax (initialize_flo.F90:620-666), t_diff = d^2/gamma."""
    from conftest import make_setup
    vis, pr = 0.012, 0.05          # a small Pr makes the diffusive limit the smaller one
    st = make_setup(tmp_path, "SC", kind="simple", my=16, dbydx=4, re=20.0, iscalon=1, pr=pr, vis=vis)
    sc = st.scalar
    assert sc["iscalon"] == 1 and abs(sc["gamma"] - vis / pr) < 1e-15 and sc["phisurf"] == 0.0
    s = st.finish(100)
    dx = st.config.dx
    phi = 100 / (st.nx * st.my * st.mz)
    dt_diff = dx * dx * (1 - phi) / (vis / pr)
    assert abs(s.dt - 0.1 * dt_diff) < 1e-15 * dt_diff   # CFL = 0.1 of write_case
    host = _abi.load_host()
    tc, tv, td = (ctypes.c_double() for _ in range(3))
    host.pib_host_time_scales(ctypes.byref(tc), ctypes.byref(tv), ctypes.byref(td))
    assert abs(td.value - 1.0 / (vis / pr)) < 1e-12   # DIA_PHYS = 1


def test_scalar_deck_limits(tmp_path):
    from conftest import make_setup
    from pureibm_ps2d_fd1d_amd import configs
    d = str(tmp_path)
    configs.simple_cubic(d, "S2", my=16, dbydx=4, iscalon=1)
    p = os.path.join(d, "S2_floparam.in")
    txt = open(p).read().replace("NSPMX = 1", "NSPMX = 2")
    open(p, "w").write(txt)
    with pytest.raises(solver.PibError) as e:
        solver.RunSetup(os.path.join(d, "S2"))
    assert e.value.code == _abi.PIB_ERR_UNSUPPORTED

