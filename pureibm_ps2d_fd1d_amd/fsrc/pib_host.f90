! pib_host.f90 -- host side of the MI355X fluid step: the reference's namelist
! input contract and the part of its run set-up that the time loop consumes.
!
! Mirrors, for a fresh start of a fixed-particle run:
!   set_default / read_nmls        initialize.F90:73-359
!   generate_configuration         initialize_flo.F90:1016-1605 ('simple',
!                                  'fcc', 'default' with <RUN>_sphr_center.inp)
!   initflo                        initialize_flo.F90:61-946 (dt, cf, umean,
!                                  ufmean_des, radii, ramp time)
!   quadgener/generbnd/dcgen       geom_init.F90:44-369, 573-586
! The set-up is split in two because maxvolfrac needs the solid mask, which
! the device computes (pib_set_particles):
!   pib_host_configure  -> geometry, point set, pib_config
!   pib_host_finish     -> mask-dependent scalars (pib_scalars)
! All entry points are bind(C) so the Fortran driver and other hosts share them.
module pib_host
  use iso_c_binding
  use pib_types
  implicit none
  private
  integer, parameter :: prcn = 8
  integer, parameter :: undefined_i = -2147483647
  real(prcn), parameter :: undefined_r = -1.0d30
  real(prcn), parameter :: one = 1.0_prcn, two = 2.0_prcn, three = 3.0_prcn, half = 0.5_prcn
  real(prcn), parameter :: zero = 0.0_prcn, four = 4.0_prcn
  ! constants.F90:92-93 (single-precision literals widened to real(8))
  real(prcn), parameter :: small_number = 1.0E-35, large_number = 1.0E+32
  real(prcn) :: pi, twopi

  ! ---- namelist variables (global_data.F90:160-336, scalar_data.F90:89) ----
  character(len=40) :: errfile, outputfile
  integer :: minunitno, maxunitno, saveitns, mbuffer
  integer :: aliasflag, maxzeroflag, revernorm, imove, iscalon, irestart, iscal_restart, &
       saveforrestart, iturbon
  logical :: xperiodic, adaptive, set_umean, set_mpg, rk3, mean_vel_to_particles, dorpr, &
       dobnd, gof_avg, write_snapshot, only_dem, include_nmis, nlcons, impose_grav, &
       include_frmeanfluid, cage_simul, fixed_moving, moving_from_fixed_bed, use_drag_law, &
       initialize_zero, debug_check, movingcv, from_post, change_frame, write_output
  character(len=11) :: rstsave, outformat, rstread
  real(prcn) :: dr, f1, f2, tstartmove, saveforres_time, lybyd, dbydx, dia_phys, hbyd, &
       widthbyd, aspect_ratio, volume_fraction
  integer :: nbody
  character(len=80) :: input_type, discr_scheme, limiter, collision_type, vel_distrib, psd_type
  real(prcn) :: rhos, cps, ks, rhof, cpf, kf
  real(prcn) :: tstart, tend, tol_ferror, tol_gran_error, wtime_maxhrs, vis, re, cfl, vfl, &
       mpg(3), cfl_max, ret, yalpha(2), flo_ang(3), archno
  integer :: mx, mxf, my, mz, foffset, ramp_frac_steps, nerr_steps, nphases, mis_hs, mis_mat, &
       min_colisions
  real(prcn) :: vol_frac1, min_part_sep, nl_slope_factor, coeff_rest, pvel_var, volfrac_rat, &
       dia_ratio, red1, sigmabydavg, phiavg, hhat, matern_treshold
  logical :: equal_momentum, ibordel
  ! scal_propt (read only when ISCALON = 1)
  integer :: nspmx
  real(prcn) :: phistream, phisurf, pr_or_sc_nu
  logical :: setphimean, sourcepresent, lumped_cap, zero_flow_in_solid, dorpr_scal
  ! postlist
  integer :: nmis, nbins, skip_num, sa_max_itr, igeometry, isa, sa_in_itr, icohesive, &
       nsmall_length, cluster_part_min, cluster_part_max
  real(prcn) :: rad_factor, power, post_error, post_small, rad_ratio, func_rad_ratio, &
       cooling_rate, temp_init, sa_error, objective_coef1, objective_coef2, &
       rad_reduction_ratio, lybyd_small
  logical :: post_no_flow_mem_alloc, rotated

  namelist /flags/ aliasflag, maxzeroflag, revernorm, imove, iscalon, irestart, iscal_restart, &
       saveforrestart, saveforres_time, xperiodic, adaptive, set_umean, set_mpg, nlcons, rk3, &
       mean_vel_to_particles, dorpr, dobnd, include_frmeanfluid, impose_grav, debug_check, &
       cage_simul, fixed_moving, gof_avg, only_dem, movingcv, iturbon, moving_from_fixed_bed, &
       from_post, change_frame, use_drag_law, initialize_zero
  namelist /files/ outputfile, rstsave, outformat, saveitns, minunitno, maxunitno, errfile, &
       rstread, write_output, write_snapshot
  namelist /part_propt/ nbody, dr, f1, f2, tstartmove, lybyd, dbydx, dia_phys, rhos, cps, ks, &
       input_type, discr_scheme, limiter, vol_frac1, min_part_sep, nphases, collision_type, &
       vel_distrib, nl_slope_factor
  namelist /floandgrid/ mx, mxf, my, mz, tstart, ramp_frac_steps, tend, nerr_steps, &
       wtime_maxhrs, saveforres_time, tol_ferror, vis, re, rhof, cpf, kf, archno, cfl, cfl_max, &
       vfl, foffset, mbuffer, mpg, ret, flo_ang, tol_gran_error
  namelist /psd/ psd_type, nphases, sigmabydavg, phiavg, equal_momentum
  namelist /gcg/ coeff_rest, pvel_var, min_colisions, mis_hs
  namelist /bidisperse/ volfrac_rat, red1, dia_ratio, yalpha
  namelist /lubtest/ hbyd
  namelist /matinput/ hhat, mis_mat, ibordel, matern_treshold
  namelist /riser/ volume_fraction, widthbyd, aspect_ratio
  namelist /scal_propt/ nspmx, phistream, setphimean, phisurf, sourcepresent, lumped_cap, &
       pr_or_sc_nu, zero_flow_in_solid, dorpr_scal
  namelist /postlist/ nmis, nbins, skip_num, post_no_flow_mem_alloc, rad_factor, power, &
       sa_max_itr, post_error, post_small, rad_ratio, func_rad_ratio, cooling_rate, temp_init, &
       sa_error, objective_coef1, objective_coef2, igeometry, isa, sa_in_itr, &
       rad_reduction_ratio, icohesive, nsmall_length, lybyd_small, cluster_part_min, &
       cluster_part_max, rotated, include_nmis

  ! ---- set-up results ----
  real(prcn), allocatable :: xc(:,:), radbdy(:), bnd(:,:), cdn(:,:)
  integer :: nbnd = 0
  real(prcn) :: dx, doml(3), char_length, uchar(3), ucharmod, quad_r
  type(pib_config), save :: cfg
  ! time scales of the drag output (initialize_flo.F90:556-574, 627)
  real(prcn) :: ts_conv = 1.d0, ts_vis = 1.d0, ts_diff = large_number
  character(len=512) :: last_err = ''
  logical :: configured = .false.

  public :: pib_host_read, pib_host_configure, pib_host_finish, pib_host_counts, &
       pib_host_get_config, pib_host_get_particles, pib_host_get_points, &
       pib_host_last_error, pib_host_quadgener, pib_host_get_nsteps_hint, pib_host_get_tol_ferror, &
       pib_host_set_option, pib_host_time_scales, pib_host_save_restart, pib_host_read_restart, &
       pib_host_get_dem_config, pib_host_get_scalar, pib_host_save_scal_restart, pib_host_read_scal_restart
contains

  subroutine calculate_constants
    pi = 4.0_prcn*atan(one)
    twopi = 2.0*pi
  end subroutine calculate_constants

  ! set_default, initialize.F90:73-271 (the variables this host reads)
  subroutine set_default
    mx = undefined_i; my = undefined_i; mz = undefined_i; mxf = undefined_i
    nphases = 1
    ramp_frac_steps = 20
    nerr_steps = 100
    wtime_maxhrs = 23.d0
    tol_ferror = 1.E-06
    tol_gran_error = 1.E-03
    aliasflag = 0; maxzeroflag = 0; revernorm = 1
    dorpr = .true.; dobnd = .true.
    imove = 0; iscalon = 0; irestart = 0; iscal_restart = undefined_i
    saveforrestart = 1; saveforres_time = 6.d0
    xperiodic = .true.; adaptive = .true.; nlcons = .true.
    set_umean = .false.; set_mpg = .false.; rk3 = .false.
    mean_vel_to_particles = .false.; include_frmeanfluid = .false.; impose_grav = .false.
    cage_simul = .false.; gof_avg = .false.; write_snapshot = .false.; fixed_moving = .false.
    only_dem = .false.; movingcv = .false.; debug_check = .true.; iturbon = 0
    moving_from_fixed_bed = .false.; from_post = .false.; change_frame = .false.
    use_drag_law = .false.; initialize_zero = .false.; include_nmis = .false.
    outputfile = "ffd.out"; rstsave = "unformatted"; outformat = "formatted"
    errfile = "ffd.err"; rstread = "unformatted"
    saveitns = 15000; write_output = .false.; minunitno = 120; maxunitno = 300
    nbody = 1; dr = 1.0; f1 = 1.6d0; f2 = 1.6d0; dia_phys = 1.d0; tstartmove = 0.00744
    lybyd = undefined_r; dbydx = undefined_r
    rhos = 1.0; cps = 447.0; ks = 0.03
    vol_frac1 = 0.1d0; volfrac_rat = 1.0; nl_slope_factor = 1.d0
    input_type = "default"; discr_scheme = "center"; limiter = "none"
    vel_distrib = "maxwellian"; collision_type = "none"; psd_type = "discrete"
    widthbyd = lybyd; aspect_ratio = one; volume_fraction = 0.05
    min_part_sep = 0.d0; dia_ratio = 1.d0
    vis = 0.012; re = 100.00; cfl = 0.5000E+00; cfl_max = cfl; vfl = 1.0
    foffset = 0; mbuffer = 0; ret = zero; mpg = 0.0
    flo_ang(1) = 0.d0; flo_ang(2) = 90.d0; flo_ang(3) = 90.d0
    coeff_rest = 1.0; pvel_var = 1.0; min_colisions = 10; mis_hs = 1
    hhat = 2.d0; mis_mat = 1; ibordel = .false.; matern_treshold = 0.9
    nspmx = 1; phistream = 1.d0; phisurf = 0.d0; sourcepresent = .false.
    pr_or_sc_nu = 0.7d0; zero_flow_in_solid = .false.; dorpr_scal = .false.; setphimean = .true.
    lumped_cap = .false.
    post_no_flow_mem_alloc = .false.; from_post = .false.
    nmis = 1; rad_factor = 2.0; power = 1.0; nbins = 200; sa_max_itr = 1
    post_error = 1d-3; post_small = 1d-3; rad_ratio = one; cooling_rate = 0.95
    func_rad_ratio = 6; temp_init = 1000; objective_coef1 = one; objective_coef2 = one
    igeometry = 0; isa = 0; sa_in_itr = 1; rad_reduction_ratio = 0.9; icohesive = 0
    nsmall_length = 1; cluster_part_min = 0; cluster_part_max = 0; lybyd_small = one
    rotated = .false.
    tstart = zero; tend = zero; rhof = one; cpf = zero; kf = zero; archno = 20.0
    sigmabydavg = zero; phiavg = zero; equal_momentum = .false.
    red1 = zero; yalpha = zero; hbyd = zero; sa_error = zero
  end subroutine set_default

  subroutine cstr(src, n, dst)
    character(kind=c_char), intent(in) :: src(*)
    integer(c_int), intent(in) :: n
    character(len=*), intent(out) :: dst
    integer :: i
    dst = ''
    do i = 1, min(int(n), len(dst))
       dst(i:i) = src(i)
    end do
  end subroutine cstr

  ! read_nmls, initialize.F90:274-359.  Every group is searched from the top of
  ! the file; FLOANDGRID, PART_PROPT and FLAGS are required, the others optional.
  function pib_host_read(path, n) bind(C, name="pib_host_read") result(st)
    character(kind=c_char), intent(in) :: path(*)
    integer(c_int), value :: n
    integer(c_int) :: st
    character(len=1024) :: fname
    integer :: u, ios
    call calculate_constants
    call set_default
    call cstr(path, n, fname)
    st = PIB_OK
    configured = .false.
    open(newunit=u, file=trim(fname), status='old', form='formatted', delim='apostrophe', iostat=ios)
    if (ios /= 0) then
       last_err = 'cannot open namelist file '//trim(fname)
       st = PIB_ERR_ARG; return
    end if
    read(u, nml=floandgrid, iostat=ios)
    if (ios /= 0) then
       last_err = 'error reading namelist FLOANDGRID'; st = PIB_ERR_ARG; close(u); return
    end if
    rewind(u)
    read(u, nml=part_propt, iostat=ios)
    if (ios /= 0) then
       last_err = 'error reading namelist PART_PROPT'; st = PIB_ERR_ARG; close(u); return
    end if
    rewind(u)
    read(u, nml=flags, iostat=ios)
    if (ios /= 0) then
       last_err = 'error reading namelist FLAGS'; st = PIB_ERR_ARG; close(u); return
    end if
    if (imove /= 1) cage_simul = .false.
    if (imove /= 1) fixed_moving = .false.
    if (set_umean) set_mpg = set_umean
    rewind(u); read(u, nml=files, iostat=ios)
    rewind(u); read(u, nml=psd, iostat=ios)
    rewind(u); read(u, nml=gcg, iostat=ios)
    if (trim(input_type) == 'lubtest') then
       rewind(u); read(u, nml=lubtest, iostat=ios)
    end if
    if (iscalon == 1) then
       rewind(u); read(u, nml=scal_propt, iostat=ios)
    end if
    rewind(u); read(u, nml=postlist, iostat=ios)
    close(u)
    if (my /= undefined_i) then
       if (mz == undefined_i) mz = my
       if (mx == undefined_i) mx = my + 1
       mxf = mx
    end if
  end function pib_host_read

  ! Harness overrides applied after reading: 1 = precision (0 FP64, 1 FP32-bulk).
  function pib_host_set_option(key, val) bind(C, name="pib_host_set_option") result(st)
    integer(c_int), value :: key
    real(c_double), value :: val
    integer(c_int) :: st
    st = PIB_OK
    select case (key)
    case (1)
       cfg%precision = int(val, c_int)
    case default
       st = PIB_ERR_ARG
    end select
  end function pib_host_set_option

  ! quadgener + generbnd + dcgen, geom_init.F90:44-369, 573-586.  Returns the
  ! point count; fills bndpts(3,npt)/cd(3,npt) when maxpts allows.
  function pib_host_quadgener(r, f2in, maxpts, bout, cout) bind(C, name="pib_host_quadgener") result(cnt)
    real(c_double), value :: r, f2in
    integer(c_int), value :: maxpts
    real(c_double), intent(out) :: bout(3,*), cout(3,*)
    integer(c_int) :: cnt
    real(prcn), allocatable :: nxb(:), nyb(:), nzb(:)
    integer :: imax, jmax, i, j, npt, cap, g
    real(prcn) :: dtheta, theta, phi, dphi, x, y, z, rr
    integer, parameter :: sx(8) = (/1, -1, 1, 1, -1, -1, 1, -1/)
    integer, parameter :: sy(8) = (/1, 1, -1, 1, -1, 1, -1, -1/)
    integer, parameter :: sz(8) = (/1, 1, 1, -1, 1, -1, -1, -1/)
    logical :: need
    call calculate_constants
    imax = nint(twopi*r/four*f2in)
    dtheta = twopi/imax/four
    cap = 0
    theta = 0.
    do i = 1, imax + 1
       jmax = nint(twopi*r*sin(theta)/4*f2in)
       cap = cap + jmax + 1
       theta = theta + dtheta
    end do
    allocate(nxb(cap), nyb(cap), nzb(cap))
    npt = 0
    theta = 0.
    do i = 1, imax + 1
       jmax = nint(twopi*r*sin(theta)/4*f2in)
       dphi = twopi/jmax/4
       phi = 0.
       do j = 1, jmax + 1
          z = r*cos(theta)/r
          if (abs(z) < 1.0d-5) z = 0.
          y = r*sin(theta)*cos(phi)/r
          if (abs(y) < 1.0d-5) y = 0.
          x = r*sin(theta)*sin(phi)/r
          if (abs(x) < 1.0d-5) x = 0.
          npt = npt + 1
          nxb(npt) = x; nyb(npt) = y; nzb(npt) = z
          phi = phi + dphi
       end do
       theta = theta + dtheta
    end do
    cnt = 0
    do g = 1, 8
       do i = 1, npt
          need = .true.
          if (sx(g) < 0 .and. nxb(i) == 0.d0) need = .false.
          if (sy(g) < 0 .and. nyb(i) == 0.d0) need = .false.
          if (sz(g) < 0 .and. nzb(i) == 0.d0) need = .false.
          if (.not. need) cycle
          cnt = cnt + 1
          if (cnt <= maxpts) then
             x = sx(g)*nxb(i); y = sy(g)*nyb(i); z = sz(g)*nzb(i)
             bout(1,cnt) = x; bout(2,cnt) = y; bout(3,cnt) = z
             rr = dsqrt(dble(x*x + y*y + z*z))
             cout(1,cnt) = dble(x)/rr; cout(2,cnt) = dble(y)/rr; cout(3,cnt) = dble(z)/rr
          end if
       end do
    end do
    deallocate(nxb, nyb, nzb)
  end function pib_host_quadgener

  ! generate_configuration ('simple', 'fcc', 'default') + the geometry part of
  ! initflo, then quadgener.  run_name selects <RUN>_sphr_center.inp.
  function pib_host_configure(run, n) bind(C, name="pib_host_configure") result(st)
    character(kind=c_char), intent(in) :: run(*)
    integer(c_int), value :: n
    integer(c_int) :: st
    character(len=1024) :: run_name
    integer :: u, ios, i, cnt, mx1
    real(prcn) :: dummyx, dummyy, dummyz, dummyr, dia
    real(prcn) :: btmp(3,1), ctmp(3,1)
    call calculate_constants
    call cstr(run, n, run_name)
    st = PIB_OK
    configured = .false.
    if (.not. xperiodic) then
       last_err = 'XPERIODIC = F (inflow/outflow) is not supported'; st = PIB_ERR_UNSUPPORTED; return
    end if
    if (aliasflag /= 0) then
       last_err = 'ALIASFLAG = 1 is not supported'; st = PIB_ERR_UNSUPPORTED; return
    end if
    if (iscalon /= 0) then
       ! one species, no source term (the GPU step's scope; scalar_init.F90:58-68, 284-350)
       if (nspmx /= 1) then
          last_err = 'NSPMX > 1: one scalar species is supported'; st = PIB_ERR_UNSUPPORTED; return
       end if
       if (sourcepresent) then
          last_err = 'SOURCEPRESENT = T (add_source) is not supported'; st = PIB_ERR_UNSUPPORTED; return
       end if
       if (.not. (pr_or_sc_nu > zero)) then
          last_err = 'PR_OR_SC_NU must be > 0'; st = PIB_ERR_ARG; return
       end if
    end if
    if (imove == 1 .and. (cage_simul .or. only_dem)) then
       last_err = 'IMOVE with CAGE_SIMUL or ONLY_DEM: run the DEM alone through pib_dem_* instead'
       st = PIB_ERR_UNSUPPORTED; return
    end if
    if (impose_grav .or. mean_vel_to_particles .or. movingcv .or. irestart /= 0 .or. only_dem) then
       last_err = 'IMPOSE_GRAV/MEAN_VEL_TO_PARTICLES/MOVINGCV/IRESTART/ONLY_DEM are not supported'
       st = PIB_ERR_UNSUPPORTED; return
    end if
    if (allocated(xc)) deallocate(xc)
    if (allocated(radbdy)) deallocate(radbdy)
    select case (trim(input_type))
    case ('simple')
       nbody = 1; nphases = 1
       char_length = dia_phys
       lybyd = (pi/(6.d0*vol_frac1))**(one/three)
       doml(2) = lybyd*dia_phys
       if (my == undefined_i) call set_my_from_lybyd
       allocate(xc(nbody,3), radbdy(nbody))
       mx1 = mx - 1
       xc(1:nbody,1) = mx1/two + one
       xc(1:nbody,2) = my/2.d0 + one
       xc(1:nbody,3) = mz/2.d0 + one
       dx = doml(2)/(my)
       doml(1) = (mx-1)*dx
       doml(3) = doml(2)
       radbdy(1:nbody) = half*(my)/lybyd
    case ('fcc')
       nbody = 4; nphases = 1
       char_length = dia_phys
       lybyd = (4.d0*pi/(6.d0*vol_frac1))**(one/three)
       if (my == undefined_i) call set_my_from_lybyd
       allocate(xc(nbody,3), radbdy(nbody))
       xc(1,1:3) = 1.d0
       xc(2,1) = my/2.d0+one; xc(2,2) = my/2.d0+one; xc(2,3) = 1.d0
       xc(3,1) = my/2.d0+one; xc(3,2) = 1.d0; xc(3,3) = my/2.d0+one
       xc(4,1) = 1.d0; xc(4,2) = my/2.d0+one; xc(4,3) = my/2.d0+one
       doml(2) = lybyd*dia_phys
       dx = doml(2)/(my)
       doml(1) = (mx-1)*dx
       doml(3) = doml(2)
       radbdy(1:nbody) = half*(my)/lybyd
    case ('default')
       open(newunit=u, file=trim(run_name)//'_sphr_center.inp', status='old', action='read', iostat=ios)
       if (ios /= 0) then
          last_err = 'cannot open '//trim(run_name)//'_sphr_center.inp'; st = PIB_ERR_ARG; return
       end if
       cnt = 0
       read(u, *, iostat=ios)
       do
          read(u, *, iostat=ios) dummyx, dummyy, dummyz, dummyr
          if (ios /= 0) exit
          cnt = cnt + 1
       end do
       nbody = cnt
       rewind(u)
       allocate(xc(nbody,3), radbdy(nbody))
       read(u, *, iostat=ios)
       do i = 1, nbody
          read(u, *, iostat=ios) dummyx, dummyy, dummyz, dummyr
          xc(i,1) = dummyx*dbydx
          xc(i,2) = dummyy*dbydx
          xc(i,3) = dummyz*dbydx
       end do
       close(u)
       nphases = 1
       char_length = dia_phys
       doml(2) = lybyd*dia_phys
       if (my == undefined_i) call set_my_from_lybyd
       do i = 1, nbody
          radbdy(i) = half*real(my,prcn)/lybyd
       end do
       dx = doml(2)/(my)
       doml(1) = (mx-1)*dx
       doml(3) = doml(2)
    case default
       last_err = 'INPUT_TYPE '//trim(input_type)//' is not supported (simple, fcc, default)'
       st = PIB_ERR_UNSUPPORTED; return
    end select
    if (nbody < 1) then
       last_err = 'no particles'; st = PIB_ERR_ARG; return
    end if
    if (mod(my,2) /= 0 .or. mod(mz,2) /= 0 .or. mx - 1 < 4) then
       last_err = 'MY, MZ must be even and MX-1 >= 4'; st = PIB_ERR_ARG; return
    end if
    ! uchar / ucharmod, initialize_flo.F90:278-280
    uchar(1:3) = ((re*vis)/char_length)*cos(pi*flo_ang(1:3)/180.d0)
    ucharmod = sqrt(uchar(1)**2.d0 + uchar(2)**2.d0 + uchar(3)**2.d0)
    ! point set, initialize_flo.F90:462-499 (one phase)
    dia = radbdy(1)*two*dx
    quad_r = 0.5*dia/doml(2)*my
    nbnd = pib_host_quadgener(quad_r, f2, 0, btmp, ctmp)
    if (allocated(bnd)) deallocate(bnd)
    if (allocated(cdn)) deallocate(cdn)
    allocate(bnd(3,nbnd), cdn(3,nbnd))
    cnt = pib_host_quadgener(quad_r, f2, nbnd, bnd, cdn)
    ! run constants
    cfg%nx = mx - 1; cfg%my = my; cfg%mz = mz
    cfg%dx = dx; cfg%vis = vis; cfg%cfl_max = cfl_max
    cfg%rk3 = merge(1, 0, rk3); cfg%adaptive = merge(1, 0, adaptive)
    select case (trim(discr_scheme))
    case ('center'); cfg%discr_scheme = PIB_SCHEME_CENTER
    case ('upwind'); cfg%discr_scheme = PIB_SCHEME_UPWIND
    case default
       last_err = 'DISCR_SCHEME '//trim(discr_scheme)//' is not supported'; st = PIB_ERR_UNSUPPORTED; return
    end select
    select case (trim(limiter))
    case ('minmod'); cfg%limiter = PIB_LIMITER_MINMOD
    case default; cfg%limiter = PIB_LIMITER_NONE
    end select
    cfg%revernorm = revernorm; cfg%dorpr = merge(1, 0, dorpr); cfg%dobnd = merge(1, 0, dobnd)
    cfg%include_frmeanfluid = merge(1, 0, include_frmeanfluid)
    cfg%set_mpg = merge(1, 0, set_mpg); cfg%soft_start = 1
    cfg%ucharmod = ucharmod; cfg%char_length = char_length; cfg%dia = dia
    cfg%dr = dr; cfg%voldom = doml(1)*doml(2)*doml(3)
    ! IMOVE = 1 moves the particles from the first step (dtibm_main_scal_flo.F90:
    ! 426-431); with FIXED_MOVING the reference waits for the fixed-bed flow
    ! to converge, which the caller decides here by setting the flag later
    cfg%move_particles = merge(1, 0, imove == 1 .and. .not. fixed_moving)
    cfg%rhos = rhos; cfg%rhof = rhof
    configured = .true.
  contains
    subroutine set_my_from_lybyd
      my = lybyd*dbydx
      if (mod(my,2) /= 0) my = my + 1
      mz = my
      mx = my + 1
      mxf = mx
    end subroutine set_my_from_lybyd
  end function pib_host_configure

  function pib_host_counts(nb, np, nx, my_, mz_) bind(C, name="pib_host_counts") result(st)
    integer(c_int), intent(out) :: nb, np, nx, my_, mz_
    integer(c_int) :: st
    nb = nbody; np = nbnd; nx = mx - 1; my_ = my; mz_ = mz
    st = merge(PIB_OK, PIB_ERR_STATE, configured)
  end function pib_host_counts

  function pib_host_get_config(c) bind(C, name="pib_host_get_config") result(st)
    type(pib_config), intent(out) :: c
    integer(c_int) :: st
    c = cfg
    st = merge(PIB_OK, PIB_ERR_STATE, configured)
  end function pib_host_get_config

  function pib_host_get_particles(xo, ro) bind(C, name="pib_host_get_particles") result(st)
    real(c_double), intent(out) :: xo(nbody,3), ro(nbody)
    integer(c_int) :: st
    st = merge(PIB_OK, PIB_ERR_STATE, configured)
    if (.not. configured) return
    xo = xc; ro = radbdy
  end function pib_host_get_particles

  function pib_host_get_points(bo, co) bind(C, name="pib_host_get_points") result(st)
    real(c_double), intent(out) :: bo(3,nbnd), co(3,nbnd)
    integer(c_int) :: st
    st = merge(PIB_OK, PIB_ERR_STATE, configured)
    if (.not. configured) return
    bo = bnd; co = cdn
  end function pib_host_get_points

  ! Mask-dependent part of initflo for a fresh fixed-bed start,
  ! initialize_flo.F90:180-207, 278-359, 447-699.
  function pib_host_finish(count_solid, s) bind(C, name="pib_host_finish") result(st)
    integer(c_long_long), value :: count_solid
    type(pib_scalars), intent(out) :: s
    integer(c_int) :: st
    real(prcn) :: maxvolfrac, vbox, volfrac, mean_volfrac, velbdy_max, umeanslip, umax_tmp
    real(prcn) :: dt_tmp_conv, dt_tmp_vis, dt_tmp_grav, dt_tmp_diff, dtcoll, dt, cf, lchar, t_vis, t_min, t_conv
    real(prcn) :: ufmean_des(3), usmean_des(3)
    integer :: idim
    real(c_double) :: zeros88(88)
    st = PIB_OK
    zeros88 = 0.d0
    s = transfer(zeros88, s)
    if (.not. configured) then
       st = PIB_ERR_STATE; last_err = 'pib_host_configure first'; return
    end if
    maxvolfrac = real(count_solid,prcn)/real(((mx-1)*my*mz),prcn)
    vbox = doml(1)*doml(2)*doml(3)
    volfrac = (nbody)*pi*(cfg%dia**3.d0)/(6.d0*vbox)
    mean_volfrac = zero + volfrac
    ! particles at rest (generate_particle_velocities with ReT = 0)
    usmean_des = zero
    do idim = 1, 3
       ufmean_des(idim) = uchar(idim)/(one-maxvolfrac)
       if (set_mpg .or. initialize_zero) then
          s%umean(idim) = zero
       else
          s%umean(idim) = (one-maxvolfrac)*ufmean_des(idim) + maxvolfrac*usmean_des(idim)
       end if
    end do
    umeanslip = sqrt(uchar(1)**2.d0 + uchar(2)**2.d0 + uchar(3)**2.d0)/(one-maxvolfrac)
    t_conv = (char_length)/(umeanslip*(one-maxvolfrac))
    t_min = t_conv
    velbdy_max = zero
    umax_tmp = max(umeanslip, velbdy_max)
    dt_tmp_conv = (dx)/(umax_tmp)
    lchar = char_length
    t_vis = (lchar*lchar)/vis
    dt_tmp_vis = large_number
    dt_tmp_diff = large_number
    if (t_vis - t_min < small_number) t_min = t_vis
    lchar = dx
    if (re < one) dt_tmp_vis = re*dbydx*dt_tmp_conv/(cfl*vfl)
    ! initscal (scalar_init.F90:73-98): gamma = vis/Pr, dt_tmp_diff = dx^2 (1 - maxvolfrac)/gamma
    if (iscalon == 1) dt_tmp_diff = (dx*dx*(one-maxvolfrac))/(vis/pr_or_sc_nu)
    dt_tmp_grav = large_number
    dtcoll = large_number
    dt = min(dt_tmp_conv, dt_tmp_vis, dt_tmp_grav, dt_tmp_diff)
    dt = cfl*dt
    dt = min(dt, dtcoll)
    cf = -1.d0/dt
    s%t = tstart
    s%dt = dt; s%cf = cf; s%cforig = cf
    s%ufmean_des = ufmean_des; s%usmean_des = usmean_des
    s%mpg = mpg
    s%maxvolfrac = maxvolfrac; s%mean_volfrac = mean_volfrac
    s%ferror = one; s%fold = zero
    s%count_solid = real(count_solid, prcn)
    s%count_fluid = real((mx-1)*my*mz, prcn) - real(count_solid, prcn)
    s%ramp_frac_time = ramp_frac_steps*ucharmod*dt/(float(mx-1)*dx)
    ts_conv = t_conv
    ts_vis = (char_length*char_length)/vis
    ts_diff = large_number
    if (iscalon == 1) ts_diff = (char_length*char_length)/(vis/pr_or_sc_nu)
  end function pib_host_finish

  ! The scalar set-up of the run (scal_propt, scalar_init.F90:58-204):
  ! out = (iscalon, gamma = vis/Pr, phisurf, phistream, Pr, iscal_restart,
  ! voldom, char_length); valid after pib_host_configure.
  function pib_host_get_scalar(out) bind(C, name="pib_host_get_scalar") result(st)
    real(c_double), intent(out) :: out(8)
    integer(c_int) :: st
    st = PIB_OK
    out = zero
    if (.not. configured) then
       st = PIB_ERR_STATE; last_err = 'pib_host_configure first'; return
    end if
    out(1) = real(iscalon, prcn)
    out(2) = vis/pr_or_sc_nu
    out(3) = phisurf
    out(4) = phistream
    out(5) = pr_or_sc_nu
    out(6) = merge(real(iscal_restart, prcn), -one, iscal_restart /= undefined_i)
    out(7) = doml(1)*doml(2)*doml(3)
    out(8) = char_length
  end function pib_host_get_scalar

  ! t_conv, t_vis, t_diff of the _norm_drag.dat time columns
  ! (boundary_condition.F90:861); valid after pib_host_finish
  function pib_host_time_scales(tc, tv, td) bind(C, name="pib_host_time_scales") result(st)
    real(c_double), intent(out) :: tc, tv, td
    integer(c_int) :: st
    tc = ts_conv
    tv = ts_vis
    td = ts_diff
    st = PIB_OK
  end function pib_host_time_scales

  ! NERR_STEPS (floandgrid): the length of the drag-error history whose mean
  ! the time loop compares with TOL_FERROR (boundary_condition.F90:828-833).
  function pib_host_get_nsteps_hint() bind(C, name="pib_host_get_nsteps_hint") result(ns)
    integer(c_int) :: ns
    ns = nerr_steps
  end function pib_host_get_nsteps_hint

  ! TOL_FERROR (floandgrid): the convergence tolerance of that mean
  ! (dtibm_main_scal_flo.F90:471-476).
  function pib_host_get_tol_ferror() bind(C, name="pib_host_get_tol_ferror") result(t)
    real(c_double) :: t
    t = tol_ferror
  end function pib_host_get_tol_ferror

  ! ---- restart files (restart_funcs.F90) ------------------------------------
  ! save_part_restart + save_unformatted (:69-128, 188-362) for the flow of a
  ! one-phase run: the count alternates 1, 2 through <RUN>_RESTART; the
  ! records of <RUN>_sphere_config.rst (and _sphere_config_<c>.rst when
  ! IMOVE = 1), <RUN>_u_<c>.rst, <RUN>_nl_<c>.rst and <RUN>_p_<c>.rst are
  ! written in the reference's order and kinds with FORM = RSTSAVE
  ! ("unformatted": sequential records); the other count's files are deleted.
  ! The granular-temperature history of moving runs is not tracked here and
  ! is written as its initial value (error 1, old 0, history 1).
  function pib_host_save_restart(run, n, s, nb, xcin, radin, velin, nx_, my2_, mz_, u, nl, p, &
       ferr, count_out) bind(C, name="pib_host_save_restart") result(st)
    character(kind=c_char), intent(in) :: run(*)
    integer(c_int), value :: n, nb, nx_, my2_, mz_
    type(pib_scalars), intent(in) :: s
    real(c_double), intent(in) :: xcin(nb, 3), radin(nb), velin(nb, 3)
    complex(c_double_complex), intent(in) :: u(nx_ + 1, my2_, mz_, 3), nl(nx_ + 1, my2_, mz_, 3)
    complex(c_double_complex), intent(in) :: p(nx_, my2_, mz_)
    real(c_double), intent(in) :: ferr(*)
    integer(c_int), intent(out) :: count_out
    integer(c_int) :: st
    character(len=1024) :: run_name
    character(len=1100) :: fname
    character(len=10) :: c1
    integer :: ru, ios, count_routine, count_tmp, iphs
    logical :: ex
    real(prcn), allocatable :: ones(:)
    st = PIB_OK
    count_out = 0
    call cstr(run, n, run_name)
    if (nb /= nbody) then
       last_err = 'save_restart: nbody differs from the configured run'; st = PIB_ERR_ARG; return
    end if
    count_routine = 0
    inquire(file=trim(run_name)//'_RESTART', exist=ex)
    if (ex) then
       open(newunit=ru, file=trim(run_name)//'_RESTART', form='formatted', status='old', iostat=ios)
       if (ios == 0) read(ru, *, iostat=ios) count_routine
       close(ru)
       if (ios /= 0) count_routine = 0
    end if
    count_routine = count_routine + 1
    if (count_routine >= 3) count_routine = 1
    write(c1, '(I1)') count_routine
    allocate(ones(max(nerr_steps, 1)))
    ones = 1.d0
    ! sphere configuration
    open(newunit=ru, file=trim(run_name)//'_sphere_config.rst', form=trim(rstsave), status='unknown', &
         iostat=ios)
    if (ios /= 0) then
       last_err = 'save_restart: cannot open the sphere configuration file'; st = PIB_ERR_ARG; return
    end if
    call write_config(ru, .false.)
    close(ru, status='keep')
    if (imove == 1) then
       open(newunit=ru, file=trim(run_name)//'_sphere_config_'//trim(c1)//'.rst', form=trim(rstsave), &
            status='unknown')
       call write_config(ru, .true.)
       close(ru, status='keep')
    end if
    ! flow
    open(newunit=ru, file=trim(run_name)//'_u_'//trim(c1)//'.rst', form=trim(rstsave), status='replace')
    write(ru) s%t, s%dt
    write(ru) s%umean(1), s%umean(2), s%umean(3)
    write(ru) s%usmean_des(1), s%usmean_des(2), s%usmean_des(3)
    write(ru) s%ufmean_des(1), s%ufmean_des(2), s%ufmean_des(3)
    write(ru) s%cf
    write(ru) nerr_steps
    write(ru) s%ferror, s%fold
    write(ru) ferr(1:nerr_steps)
    do iphs = 1, nphases
       write(ru) s%ferror, s%fold
       if (imove == 1) write(ru) 1.d0, 0.d0
    end do
    do iphs = 1, nphases
       write(ru) ferr(1:nerr_steps)
       if (imove == 1) write(ru) ones(1:nerr_steps)
    end do
    write(ru) u
    close(ru, status='keep')
    open(newunit=ru, file=trim(run_name)//'_nl_'//trim(c1)//'.rst', form=trim(rstsave), status='replace')
    write(ru) nl
    close(ru, status='keep')
    open(newunit=ru, file=trim(run_name)//'_p_'//trim(c1)//'.rst', form=trim(rstsave), status='replace')
    write(ru) s%mpg(1), s%mpg(2), s%mpg(3)
    write(ru) p
    close(ru, status='keep')
    open(newunit=ru, file=trim(run_name)//'_RESTART', form='formatted', status='replace')
    write(ru, *) c1
    close(ru, status='keep')
    ! delete_restart_files of the other count (:364-395)
    count_tmp = 3 - count_routine
    write(c1, '(I1)') count_tmp
    call remove_file(trim(run_name)//'_u_'//trim(c1)//'.rst')
    call remove_file(trim(run_name)//'_nl_'//trim(c1)//'.rst')
    call remove_file(trim(run_name)//'_p_'//trim(c1)//'.rst')
    if (imove == 1) call remove_file(trim(run_name)//'_sphere_config_'//trim(c1)//'.rst')
    count_out = count_routine
  contains
    subroutine write_config(unit, moving)
      integer, intent(in) :: unit
      logical, intent(in) :: moving
      integer :: k
      write(unit) nbody
      if (trim(input_type) == 'random' .or. trim(input_type) == 'default') then
         write(unit) nphases
         do k = 1, nphases
            write(unit) nbody
         end do
         do k = 1, nphases
            write(unit) cfg%dia
         end do
         write(unit) doml(1:3)
         if (moving) write(unit) (cfg%move_particles /= 0)
      end if
      write(unit) char_length
      write(unit) xcin(1:nbody, 1:3)
      write(unit) radin(1:nbody)
      write(unit) velin(1:nbody, 1:3)
      if (moving) then
         write(unit) s%frame_vel(1:3)
         write(unit) s%frame_pos(1:3)
      end if
    end subroutine write_config
  end function pib_host_save_restart

  ! The scalar's restart file <RUN>_scal_<c>.rst (restart_funcs.F90:315-333,
  ! 384-387, 572-596), serial records in the reference's order: nu_error,
  ! nu_old; nu_error_array(1:nerr_steps); phif; nlphif(1:nx+1); phisurfall;
  ! phirmean, heat_ratio, heat_ratio_old.  count is the flow files' count
  ! (pib_host_save_restart's count_out); the other count's scalar file is
  ! deleted.  nu3 = (nu_error, nu_old, phirmean), hr2 = (heat_ratio,
  ! heat_ratio_old), nuerr(1:nerr_steps).
  function pib_host_save_scal_restart(run, n, count, nx_, my2_, mz_, nb, phif, nlphif, phisurfall, &
       nu3, hr2, nuerr) bind(C, name="pib_host_save_scal_restart") result(st)
    character(kind=c_char), intent(in) :: run(*)
    integer(c_int), value :: n, count, nx_, my2_, mz_, nb
    complex(c_double_complex), intent(in) :: phif(nx_ + 1, my2_, mz_), nlphif(nx_ + 1, my2_, mz_)
    real(c_double), intent(in) :: phisurfall(*), nu3(3), hr2(2), nuerr(*)
    integer(c_int) :: st
    character(len=1024) :: run_name
    character(len=10) :: c1
    integer :: ru, ios
    st = PIB_OK
    call cstr(run, n, run_name)
    if (count /= 1 .and. count /= 2) then
       last_err = 'save_scal_restart: count must be 1 or 2'; st = PIB_ERR_ARG; return
    end if
    write(c1, '(I1)') count
    open(newunit=ru, file=trim(run_name)//'_scal_'//trim(c1)//'.rst', form=trim(rstsave), status='replace', &
         iostat=ios)
    if (ios /= 0) then
       last_err = 'save_scal_restart: cannot open the scalar restart file'; st = PIB_ERR_ARG; return
    end if
    write(ru) nu3(1), nu3(2)
    write(ru) nuerr(1:nerr_steps)
    write(ru) phif
    write(ru) nlphif
    write(ru) phisurfall(1:nb)
    write(ru) nu3(3), hr2(1), hr2(2)
    close(ru, status='keep')
    write(c1, '(I1)') 3 - count
    call remove_file(trim(run_name)//'_scal_'//trim(c1)//'.rst')
  end function pib_host_save_scal_restart

  ! read_unformatted's scalar part (restart_funcs.F90:572-596) for count c.
  function pib_host_read_scal_restart(run, n, count, nx_, my2_, mz_, nb, phif, nlphif, phisurfall, &
       nu3, hr2, nuerr) bind(C, name="pib_host_read_scal_restart") result(st)
    character(kind=c_char), intent(in) :: run(*)
    integer(c_int), value :: n, count, nx_, my2_, mz_, nb
    complex(c_double_complex), intent(out) :: phif(nx_ + 1, my2_, mz_), nlphif(nx_ + 1, my2_, mz_)
    real(c_double), intent(out) :: phisurfall(*), nu3(3), hr2(2), nuerr(*)
    integer(c_int) :: st
    character(len=1024) :: run_name
    character(len=10) :: c1
    integer :: ru, ios
    st = PIB_OK
    call cstr(run, n, run_name)
    write(c1, '(I1)') count
    open(newunit=ru, file=trim(run_name)//'_scal_'//trim(c1)//'.rst', form=trim(rstsave), status='old', &
         action='read', iostat=ios)
    if (ios /= 0) then
       last_err = 'read_scal_restart: cannot open '//trim(run_name)//'_scal_'//trim(c1)//'.rst'
       st = PIB_ERR_ARG; return
    end if
    read(ru, iostat=ios) nu3(1), nu3(2)
    if (ios == 0) read(ru, iostat=ios) nuerr(1:nerr_steps)
    if (ios == 0) read(ru, iostat=ios) phif
    if (ios == 0) read(ru, iostat=ios) nlphif
    if (ios == 0) read(ru, iostat=ios) phisurfall(1:nb)
    if (ios == 0) read(ru, iostat=ios) nu3(3), hr2(1), hr2(2)
    close(ru)
    if (ios /= 0) then
       last_err = 'read_scal_restart: short or mismatched scalar restart file'; st = PIB_ERR_ARG
    end if
  end function pib_host_read_scal_restart

  subroutine remove_file(name)
    character(len=*), intent(in) :: name
    integer :: ru, ios
    logical :: ex
    inquire(file=name, exist=ex)
    if (.not. ex) return
    open(newunit=ru, file=name, status='old', iostat=ios)
    if (ios == 0) close(ru, status='delete')
  end subroutine remove_file

  ! read_unformatted (restart_funcs.F90:431-616) for the flow of a one-phase
  ! run, the count taken from <RUN>_RESTART; also reads the sphere
  ! configuration of <RUN>_sphere_config.rst (of <RUN>_sphere_config_<c>.rst,
  ! with the frame velocity and position, when IMOVE = 1).  Scalars read:
  ! t (TSTART), dt, umean, usmean_des, ufmean_des, cf (and cforig), ferror,
  ! fold, mpg; the
  ! ferror history keeps the first min(nerr_steps of the file, NERR_STEPS)
  ! entries and 1 after them, as the reference does.
  function pib_host_read_restart(run, n, s, nb, xco, rado, velo, nx_, my2_, mz_, u, nl, p, ferr, &
       count_out) bind(C, name="pib_host_read_restart") result(st)
    character(kind=c_char), intent(in) :: run(*)
    integer(c_int), value :: n, nb, nx_, my2_, mz_
    type(pib_scalars), intent(inout) :: s
    real(c_double), intent(out) :: xco(nb, 3), rado(nb), velo(nb, 3)
    complex(c_double_complex), intent(out) :: u(nx_ + 1, my2_, mz_, 3), nl(nx_ + 1, my2_, mz_, 3)
    complex(c_double_complex), intent(out) :: p(nx_, my2_, mz_)
    real(c_double), intent(out) :: ferr(*)
    integer(c_int), intent(out) :: count_out
    integer(c_int) :: st
    character(len=1024) :: run_name
    character(len=10) :: c1
    integer :: ru, ios, count_routine, nerr_tmp, iphs, nb_file, nph, k, itmp
    real(prcn) :: dtmp, d3(3), e1, e2
    logical :: lmove
    real(prcn), allocatable :: skip(:)
    st = PIB_OK
    count_out = 0
    call cstr(run, n, run_name)
    open(newunit=ru, file=trim(run_name)//'_RESTART', form='formatted', status='old', iostat=ios)
    if (ios /= 0) then
       last_err = 'read_restart: no '//trim(run_name)//'_RESTART file'; st = PIB_ERR_ARG; return
    end if
    read(ru, *, iostat=ios) count_routine
    close(ru)
    if (ios /= 0 .or. count_routine < 1 .or. count_routine > 2) then
       last_err = 'read_restart: bad count in the _RESTART file'; st = PIB_ERR_ARG; return
    end if
    write(c1, '(I1)') count_routine
    ! sphere configuration
    open(newunit=ru, file=trim(run_name)//'_sphere_config.rst', form=trim(rstsave), status='old', iostat=ios)
    if (ios /= 0) then
       last_err = 'read_restart: no sphere configuration file'; st = PIB_ERR_ARG; return
    end if
    read(ru) nb_file
    if (nb_file /= nb) then
       close(ru)
       last_err = 'read_restart: nbody of the file differs'; st = PIB_ERR_ARG; return
    end if
    if (trim(input_type) == 'random' .or. trim(input_type) == 'default') then
       read(ru) nph
       do k = 1, nph
          read(ru) itmp
       end do
       do k = 1, nph
          read(ru) dtmp
       end do
       read(ru) d3
    end if
    read(ru) dtmp
    read(ru) xco
    read(ru) rado
    read(ru) velo
    close(ru)
    ! a moving run keeps the centres, velocities and frame of the count in
    ! <RUN>_sphere_config_<c>.rst (save_unformatted :229-251)
    if (imove == 1) then
       open(newunit=ru, file=trim(run_name)//'_sphere_config_'//trim(c1)//'.rst', form=trim(rstsave), &
            status='old', iostat=ios)
       if (ios == 0) then
          read(ru) nb_file
          if (trim(input_type) == 'random' .or. trim(input_type) == 'default') then
             read(ru) nph
             do k = 1, nph
                read(ru) itmp
             end do
             do k = 1, nph
                read(ru) dtmp
             end do
             read(ru) d3
             read(ru) lmove
          end if
          read(ru) dtmp
          read(ru) xco
          read(ru) rado
          read(ru) velo
          read(ru) s%frame_vel(1:3)
          read(ru) s%frame_pos(1:3)
          close(ru)
       end if
    end if
    ! flow
    open(newunit=ru, file=trim(run_name)//'_u_'//trim(c1)//'.rst', form=trim(rstsave), status='old', iostat=ios)
    if (ios /= 0) then
       last_err = 'read_restart: no u file for the count'; st = PIB_ERR_ARG; return
    end if
    read(ru) s%t, s%dt
    read(ru) s%umean(1), s%umean(2), s%umean(3)
    read(ru) s%usmean_des(1), s%usmean_des(2), s%usmean_des(3)
    read(ru) s%ufmean_des(1), s%ufmean_des(2), s%ufmean_des(3)
    read(ru) s%cf
    s%cforig = s%cf
    read(ru) nerr_tmp
    read(ru) s%ferror, s%fold
    allocate(skip(max(nerr_tmp, nerr_steps, 1)))
    ferr(1:nerr_steps) = 1.d0
    if (nerr_tmp > nerr_steps) then
       read(ru) ferr(1:nerr_steps)
    else
       read(ru) ferr(1:nerr_tmp)
    end if
    do iphs = 1, nphases
       read(ru) e1, e2
       if (imove == 1 .and. .not. moving_from_fixed_bed) read(ru) e1, e2
    end do
    do iphs = 1, nphases
       if (nerr_tmp > nerr_steps) then
          read(ru) skip(1:nerr_steps)
       else
          read(ru) skip(1:nerr_tmp)
       end if
       if (imove == 1 .and. .not. moving_from_fixed_bed) then
          if (nerr_tmp > nerr_steps) then
             read(ru) skip(1:nerr_steps)
          else
             read(ru) skip(1:nerr_tmp)
          end if
       end if
    end do
    read(ru) u
    close(ru)
    open(newunit=ru, file=trim(run_name)//'_nl_'//trim(c1)//'.rst', form=trim(rstsave), status='old', iostat=ios)
    if (ios /= 0) then
       last_err = 'read_restart: no nl file for the count'; st = PIB_ERR_ARG; return
    end if
    read(ru) nl
    close(ru)
    open(newunit=ru, file=trim(run_name)//'_p_'//trim(c1)//'.rst', form=trim(rstsave), status='old', iostat=ios)
    if (ios /= 0) then
       last_err = 'read_restart: no p file for the count'; st = PIB_ERR_ARG; return
    end if
    read(ru) s%mpg(1), s%mpg(2), s%mpg(3)
    read(ru) p
    close(ru)
    count_out = count_routine
  end function pib_host_read_restart

  ! The DEM constants of this run, as des_time_march(.TRUE.) takes them from
  ! the run (collision_mod.F90:353-390, 1203-1316; initialize_flo.F90:630-635):
  ! DOML, FOFFSET, RHOS, RHOF, DIA_PHYS, COEFF_REST; KN left at the reference's.
  function pib_host_get_dem_config(dc) bind(C, name="pib_host_get_dem_config") result(st)
    type(pib_dem_config), intent(out) :: dc
    integer(c_int) :: st
    st = PIB_OK
    if (.not. configured) then
       last_err = 'configure the run first'; st = PIB_ERR_STATE; return
    end if
    dc%nbody = nbody
    dc%only_dem = merge(1, 0, only_dem)
    dc%reserved_i = 0
    dc%dx = dx; dc%dy = dx; dc%dz = dx
    dc%xlength = doml(1); dc%ylength = doml(2); dc%zlength = doml(3)
    dc%foffset = real(foffset, prcn)
    dc%rhos = rhos; dc%rhof = rhof
    dc%dia_phys = dia_phys
    dc%coeff_rest = coeff_rest
    dc%kn = zero
    dc%reserved_d = zero
  end function pib_host_get_dem_config

  ! Byte sizes of the bind(C) mirrors in pib_types, so that callers can check
  ! them against the C declarations.
  function pib_host_struct_sizes(sizes) bind(C, name="pib_host_struct_sizes") result(st)
    integer(c_int), intent(out) :: sizes(4)
    integer(c_int) :: st
    type(pib_config) :: a
    type(pib_scalars) :: b
    type(pib_dem_config) :: c
    type(pib_dem_info) :: d
    sizes(1) = int(c_sizeof(a), c_int)
    sizes(2) = int(c_sizeof(b), c_int)
    sizes(3) = int(c_sizeof(c), c_int)
    sizes(4) = int(c_sizeof(d), c_int)
    st = PIB_OK
  end function pib_host_struct_sizes

  function pib_host_last_error(buf, n) bind(C, name="pib_host_last_error") result(st)
    character(kind=c_char), intent(out) :: buf(*)
    integer(c_int), value :: n
    integer(c_int) :: st
    integer :: i, l
    l = min(len_trim(last_err), int(n) - 1)
    do i = 1, l
       buf(i) = last_err(i:i)
    end do
    buf(l+1) = c_null_char
    st = l
  end function pib_host_last_error
end module pib_host
