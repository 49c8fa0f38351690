! lbfast_host.f90 -- Fortran host of the gfx950 LB time step.
!
! The host side of LBFAST.f90 over the C-ABI of include/lbfast_hip.h:
!  * the &simulation and &fluid namelists of an input deck (LBFAST.f90:31-60),
!    read from the file named on the command line (default input.dat);
!  * the compile-time switches of lbfast_defines.h (the reference's defines.h);
!  * derived parameters (LBFAST.f90:212-219) and isfluid classification
!    (LBFAST.f90:258-328), optionally from isfluid.raw;
!  * the initial condition of init_conditions_module.f90:290-420 (Taylor-
!    Green, droplets, Poiseuille), or rho/u/v/w/phi.raw (lreadinit);
!  * the time loop in FLIP/FLOP pairs (integrator_module.f90:266-827), with
!    the statistics.f90 diagnostics every stamp_term steps: Taylor-Green
!    kinetic energy against exp(-6 nu k^2 t), the Laplace pressure jump
!    sigma_eff = (p_in - p_out) R / 2, the Lamb oscillation record (lamb.dat),
!    global mass;
!  * the end-of-run benchmark comparisons (statistics.f90:688-1114): the
!    one- and two-fluid Poiseuille profiles against the analytic ones, the
!    Lamb oscillation period against the viscous-corrected theory;
!  * performance as LBFAST.f90:500 (MLUPS); frames every stamp steps as RAW
!    (lraw, with XDMF sidecars under DOXDMF) and VTK image data (lvtk), 2-D
!    plane frames every stamp2D steps (nplanes, ndir, npoint), and the restart
!    file (lwriterestart / lrestart), in the reference's formats.
#include "lbfast_defines.h"
program lbfast_host
  use iso_c_binding
  use lbfast_hip
  implicit none
  integer, parameter :: db = kind(1.0d0)
  real(db), parameter :: pi_greek = 3.14159265358979323846264338327950288_db, cssq = 1.0_db / 3.0_db

  ! ---- &simulation (LBFAST.f90:31-37; types and defaults of vars_module.f90) --------------
  integer :: nsteps = 0, stamp = 0, stamp2D = 0, stepskip = 1, lx = 0, ly = 0, lz = 0
  integer :: pbc_x = 1, pbc_y = 1, pbc_z = 1, openbc = 0, iprobe = 0, jprobe = 0, kprobe = 0
  integer :: nplanes = 0, stamp_term = huge(1), every_time_check = 1000
  integer :: openbc_type_x(2) = 0, openbc_type_y(2) = 0, openbc_type_z(2) = 0
  real(db) :: openbc_press_x(2) = 0, openbc_press_y(2) = 0, openbc_press_z(2) = 0
  real(db) :: openbc_u_x(2) = 0, openbc_u_y(2) = 0, openbc_u_z(2) = 0
  real(db) :: openbc_v_x(2) = 0, openbc_v_y(2) = 0, openbc_v_z(2) = 0
  real(db) :: openbc_w_x(2) = 0, openbc_w_y(2) = 0, openbc_w_z(2) = 0
  real(db) :: time_limit = -1
  logical :: lreadisfluid = .false., lreadinit = .false., lprint = .false., lvtk = .false., lraw = .false.
  logical :: lrestart = .false., lasync = .false., lwriterestart = .false., lweakscaling = .false.
  namelist /simulation/ nsteps, stamp, stamp2D, stepskip, lreadisfluid, &
    lreadinit, lx, ly, lz, lprint, lvtk, lraw, lrestart, pbc_x, pbc_y, pbc_z, openbc, lasync, &
    iprobe, jprobe, kprobe, openbc_type_x, openbc_type_y, openbc_type_z, &
    openbc_press_x, openbc_press_y, openbc_press_z, &
    openbc_u_x, openbc_u_y, openbc_u_z, openbc_v_x, openbc_v_y, openbc_v_z, &
    openbc_w_x, openbc_w_y, openbc_w_z, &
    nplanes, stamp_term, time_limit, every_time_check, lwriterestart, lweakscaling

  ! ---- &fluid (LBFAST.f90:39-59); values the reference leaves unset start at zero ----------
  real(db) :: fx = 0, fy = 0, fz = 0, visc1 = 0, radius = 0, width = 0, center(3) = 0
  real(db) :: tau_diff = 0, sigma = 0, wettab_r = 0, wettab_b = 0, visc2 = 0
  real(db) :: rho_r = 1, rho_b = 1, uwall = 0, A_rep = 0, cosOppT = 0, q_th = 0
  integer :: npoint(64) = 0, ndir(64) = 0
  namelist /fluid/ fx, fy, fz, visc1, radius, width, center, &
#ifdef TWOCOMPONENT
    tau_diff, sigma, wettab_r, wettab_b, visc2, &
#ifdef DENSRATIO
    rho_r, rho_b, &
#endif
#endif
#ifdef REPULSIVE_FLUX
    A_rep, cosOppT, q_th, &
#endif
    uwall, npoint, ndir

  type(lbf_params), target :: p
  type(c_ptr) :: ctx
  character(len=256) :: infile
  integer :: rc, step, i, nargs, iframe, iframe2d, ntot, chunk, nxt
  integer(c_int8_t), allocatable, target :: isf(:, :, :)
  real(db), allocatable, target :: fld(:, :, :, :), buf(:, :, :), buf2(:, :, :)
  real(db) :: red(8), e0, tend, mlups
  integer(8) :: c0, c1, crate
  logical :: two, stop_early, lexist
  integer :: pdec(3)

  interface
    integer(c_int) function c_mkdir(path, mode) bind(C, name="mkdir")
      import
      character(kind=c_char), intent(in) :: path(*)
      integer(c_int), value :: mode
    end function
  end interface

  ! command line (LBFAST.f90:80-156): [input], or [proc_x] [proc_y] [proc_z]
  ! [input] as the MPI build takes it; one process drives one GPU, so the
  ! decomposition must be 1 1 1
  nargs = command_argument_count()
  if (nargs == 1) then
    call get_command_argument(1, infile)
  else if (nargs == 4) then
    do i = 1, 3
      call get_command_argument(i, infile)
      read (infile, *) pdec(i)
    end do
    if (product(pdec) /= 1) then
      write (6, '(a,3i6)') 'lbfast_host: one process drives one GPU; decomposition must be 1 1 1, got', pdec
      stop 2
    end if
    call get_command_argument(4, infile)
  else
    write (6, '(a)') 'the command line should be'
    write (6, '(a)') '[executable] [input]   or   [executable] 1 1 1 [input]'
    stop 2
  end if
#ifdef DOBENCHMARK
  ! benchmark defaults, overridden by the input deck (LBFAST.f90:159-191)
  nsteps = 100; stamp = 100; stamp2D = 10; stepskip = 1
  lx = MYSIDE; ly = MYSIDE; lz = MYSIDE
  radius = lx / 4
  uwall = 0; lprint = .true.; lvtk = .false.; lraw = .true.; lasync = .false.; lrestart = .false.
  pbc_x = 1; pbc_y = 1; pbc_z = 1; fx = 0; fy = 0; fz = 0
#endif
  inquire (file=trim(infile), exist=lexist)
  if (.not. lexist) then
    write (6, '(3a)') 'ERROR: file ', trim(infile), ' does not exist!'
    stop 2
  end if
  open (unit=24, file=trim(infile), status='old', action='read')
  read (24, nml=simulation)
  read (24, nml=fluid)
  close (24)

  ! ---- parameter block --------------------------------------------------------------
  call lbf_params_default(p)
  p%nx = lx; p%ny = ly; p%nz = lz
  p%lattice = LATTICE
  p%highorder = 0
#ifdef HIGHORDER
  if (LATTICE == 27) p%highorder = 1
#endif
  p%twocomponent = 0; p%densratio = 0; p%interface_incomp = 0
#ifdef TWOCOMPONENT
  p%twocomponent = 1
#endif
#ifdef DENSRATIO
  p%densratio = 1
#endif
#ifdef INTERFACE_INCOMP
  p%interface_incomp = 1
#endif
  p%internal_obstacles = 0
#ifdef INTERNAL_OBSTACLES
  p%internal_obstacles = 1
#endif
  p%buoyancy_forcing = 0
#ifdef BUOYANCY_FORCING
  p%buoyancy_forcing = 1
#endif
  p%bounce_back = 0; p%smagorinsky = 0; p%repulsive_flux = 0; p%bcphiflux = 0
#ifdef BOUNCE_BACK
  p%bounce_back = 1
#endif
#ifdef SMAGORINSKI
  p%smagorinsky = 1
#endif
#ifdef REPULSIVE_FLUX
  p%repulsive_flux = 1
#endif
#ifdef BCPHIFLUX
  p%bcphiflux = 1
#endif
  p%pbc = [pbc_x, pbc_y, pbc_z]
#if defined(STORE_FP16)
  p%precision = LBF_PREC_FP16S
#elif defined(STORE_BF16)
  p%precision = LBF_PREC_BF16S
#elif defined(MIXEDPRC)
  p%precision = LBF_PREC_MIXED
#elif PRC == 4
  p%precision = LBF_PREC_FP32
#else
  p%precision = LBF_PREC_FP64
#endif
  p%fx = fx; p%fy = fy; p%fz = fz
  p%visc1 = visc1; p%visc2 = visc2; p%rho_r = rho_r; p%rho_b = rho_b
  p%tau_diff = tau_diff; p%sigma = sigma; p%width = width; p%wettab_r = wettab_r
  p%a_rep = A_rep; p%q_th = q_th; p%cos_opp_t = cosOppT
  ! open boundaries (boundary_cds_module.f90:57-1664): faces x(1) x(2) y(1) y(2) z(1) z(2)
  p%openbc = openbc
  p%openbc_type = [openbc_type_x, openbc_type_y, openbc_type_z]
  p%openbc_press = [openbc_press_x, openbc_press_y, openbc_press_z]
  p%openbc_u = [openbc_u_x, openbc_u_y, openbc_u_z]
  p%openbc_v = [openbc_v_x, openbc_v_y, openbc_v_z]
  p%openbc_w = [openbc_w_x, openbc_w_y, openbc_w_z]
  rc = lbf_params_derive(p)
  call check(rc, 'lbf_params_derive')
  two = p%twocomponent /= 0
  write (6, '(a,3i6,a,i3,a,i2,a,i2,a,i2)') 'lbfast_host: lx ly lz', lx, ly, lz, '  lattice', p%lattice, &
    '  highorder', p%highorder, '  twocomponent', p%twocomponent, '  precision', p%precision
  write (6, '(a,7i2,i4)') 'lbfast_host: switches densratio incomp obstacles buoyancy bounce_back smagorinsky'// &
    ' repulsive pbc', p%densratio, p%interface_incomp, p%internal_obstacles, p%buoyancy_forcing, &
    p%bounce_back, p%smagorinsky, p%repulsive_flux, p%pbc(1) * 100 + p%pbc(2) * 10 + p%pbc(3)
  write (6, '(a,i2)') 'lbfast_host: bcphiflux', p%bcphiflux
  write (6, '(a,i2,a,6i2)') 'lbfast_host: openbc', p%openbc, '  types', p%openbc_type
  write (6, '(a,5g16.8)') 'lbfast_host: tau1 omega tau2 beta kapp', p%tau1, p%omega, p%tau2, p%beta, p%kapp
  flush (6)

  ! ---- isfluid (LBFAST.f90:258-328) -----------------------------------------------------
  allocate (isf(lx, ly, lz), buf(lx, ly, lz))
  isf = 1
  if (lreadisfluid) then
    call read_raw('isfluid.raw', buf)
    isf = int(buf, c_int8_t)
  end if
  rc = lbf_classify_isfluid(p, isf, isf)
  call check(rc, 'lbf_classify_isfluid')
  ! output directory of print_module.f90:326-335, made before the device is touched
  if (lprint .or. lwriterestart .or. lrestart) rc = c_mkdir('output'//c_null_char, int(o'755', c_int))

  rc = lbf_create(p, ctx)
  if (rc /= 0) then
    write (6, '(a,a)') 'lbfast_host: lbf_create failed: ', trim(lbf_error_message(c_null_ptr))
    stop 2
  end if
  if (p%internal_obstacles /= 0) then
    rc = lbf_upload_isfluid(ctx, isf)
    call check(rc, 'lbf_upload_isfluid')
  end if

  ! ---- initial condition (LBFAST.f90:336-341, integrator_module.f90:166-176) -------------
  ! the generated condition first; lreadinit then replaces P, u, v, w and phi
  ! and keeps the generated second moments, as driver_read_init_raw does;
  ! lrestart replaces the ten moments and phi
  allocate (fld(lx, ly, lz, 11))
  iframe = 0; iframe2d = 0
  call initial_fields(fld)
  if (lreadinit) then
    call read_raw('rho.raw', fld(:, :, :, 1))
    call read_raw('u.raw', fld(:, :, :, 2))
    call read_raw('v.raw', fld(:, :, :, 3))
    call read_raw('w.raw', fld(:, :, :, 4))
    if (two) call read_raw('phi.raw', fld(:, :, :, 11))
  end if
  if (lrestart) call read_restart(fld, iframe, iframe2d)
  do i = 1, 10
    rc = lbf_upload_field(ctx, int(i, c_int), c_loc(fld(1, 1, 1, i)), LBF_DTYPE_F64)
    call check(rc, 'lbf_upload_field')
  end do
  if (two) then
    rc = lbf_upload_field(ctx, LBF_FIELD_PHI, c_loc(fld(1, 1, 1, 11)), LBF_DTYPE_F64)
    call check(rc, 'lbf_upload_field phi')
  end if
  rc = lbf_init_state(ctx)
  call check(rc, 'lbf_init_state')
  deallocate (fld)

  if (lprint .and. lraw) call write_isfluid_frame()
  if (lprint .and. (lraw .or. lvtk) .and. .not. lrestart) call write_frame(iframe)
  if (lprint .and. nplanes > 0 .and. .not. lrestart) call write_frame_2d(iframe2d)

  ! ---- time loop (integrator_module.f90:266-827) ------------------------------------------
  ! The reference advances FLIP/FLOP pairs: 2*ceil(nsteps/2) steps.  Steps
  ! are enqueued in chunks that end on the next step at which output,
  ! statistics or the wall-clock check is due.
  e0 = kinetic_energy()
#if defined(LAMBTEST) && defined(TWOCOMPONENT)
  ! open_lamb (statistics.f90:277-383, called at integrator_module.f90:204)
  open (unit=142, file='lamb.dat', action='write', status='replace')
#endif
  call report(0)
  ntot = 2 * ((max(nsteps, 0) + 1) / 2)
  stop_early = .false.
  call system_clock(c0, crate)
  step = 0
  do while (step < ntot)
    nxt = ntot
    if (stamp > 0 .and. lprint) nxt = min(nxt, (step / stamp + 1) * stamp)
    if (stamp2D > 0 .and. lprint .and. nplanes > 0) nxt = min(nxt, (step / stamp2D + 1) * stamp2D)
    if (stamp_term > 0) nxt = min(nxt, (step / stamp_term + 1) * stamp_term)
    if (time_limit > 0 .and. every_time_check > 0) nxt = min(nxt, (step / every_time_check + 1) * every_time_check)
    chunk = nxt - step
    rc = lbf_step(ctx, int(chunk, c_int32_t))
    call check(rc, 'lbf_step')
    step = nxt
    if (lprint .and. stamp > 0) then
      if (mod(step, stamp) == 0) then
        iframe = iframe + 1
        if (lraw .or. lvtk) call write_frame(iframe)
      end if
    end if
    if (lprint .and. stamp2D > 0 .and. nplanes > 0) then
      if (mod(step, stamp2D) == 0) then
        iframe2d = iframe2d + 1
        call write_frame_2d(iframe2d)
      end if
    end if
    if (stamp_term > 0) then
      if (mod(step, stamp_term) == 0) call report(step)
    end if
    if (time_limit > 0 .and. every_time_check > 0) then
      if (mod(step, every_time_check) == 0) then
        rc = lbf_synchronize(ctx)
        call system_clock(c1)
        if (real(c1 - c0, db) / real(crate, db) > time_limit) then
          write (6, '(a,f16.2,a,i16,a,f16.2,a)') 'Time limit ', time_limit, ' sec reached at step ', step, &
            ' after ', real(c1 - c0, db) / real(crate, db), ' sec'
          lwriterestart = .true.
          stop_early = .true.
          exit
        end if
      end if
    end if
  end do
  rc = lbf_synchronize(ctx)
  call check(rc, 'lbf_synchronize')
  call system_clock(c1)
  tend = real(c1 - c0, db) / real(crate, db)
  ! LBFAST.f90:495-505: lattice updates per second over the run
  mlups = real(lx, db) * real(ly, db) * real(lz, db) * real(step, db) / max(tend, 1.0e-9_db) / 1.0e6_db
  write (6, '(a,i10,a,f12.4,a,f14.2)') 'lbfast_host: steps', step, '  seconds', tend, '  MLUPS', mlups
  if (lwriterestart) call write_restart(iframe, iframe2d)
  ! compare_benchmark (statistics.f90:688-1114, called at LBFAST.f90:530)
#if defined(LAMBTEST) && defined(TWOCOMPONENT)
  close (142)
  call lamb_period()
#endif
#if defined(POISEUILLE) || (defined(TWOPOISEUILLE) && defined(TWOCOMPONENT))
  call poiseuille_profile()
#endif
  call lbf_destroy(ctx)
  write (6, '(a)') 'lbfast_host: done'

contains

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= 0) then
      write (6, '(a,a,a,i3,a,a)') 'lbfast_host: ', what, ' -> ', code, ': ', trim(lbf_error_message(ctx))
      stop 2
    end if
  end subroutine

  ! fcut_tanh of vars_module.f90:1306-1329
  pure real(db) function fcut_tanh(r, r0, w)
    real(db), intent(in) :: r, r0, w
    fcut_tanh = 0.5_db * (1.0_db - tanh((r - r0) / (0.5_db * w)))
  end function

  ! minimum-image distance along one periodic axis
  ! (pbc_images, init_conditions_module.f90:591-599)
  pure real(db) function image(d, l)
    real(db), intent(in) :: d
    integer, intent(in) :: l
    image = d - real(l, db) * nint(d / real(l, db))
  end function

  ! init_conditions_module.f90:290-420
  subroutine initial_fields(f)
    real(db), intent(out) :: f(:, :, :, :)
    real(db) :: x, y, z, u, v, w, press, phi, rhophi, d(3), dist, hp, xc, cb, cl, cr, wl, wr
    integer :: gi, gj, gk
    f = 0
    phi = 0
#if defined(POISEUILLE) || defined(TWOPOISEUILLE)
#ifdef BOUNCE_BACK
    hp = 0.5_db * real(lx - 2, db)
#else
    hp = 0.5_db * real(lx - 1, db)
#endif
    xc = 0.5_db * real(lx + 1, db)
#endif
#ifdef TWOPOISEUILLE
    ! two-layer channel profile (init_conditions_module.f90:256-275)
#ifdef DENSRATIO
    cb = fz * hp * hp * (rho_r + rho_b) / (2.0_db * (rho_r * visc1 + rho_b * visc2))
    cl = fz * hp * rho_b * (visc1 - visc2) / (2.0_db * visc1 * (rho_r * visc1 + rho_b * visc2))
    cr = fz * hp * rho_r * (visc1 - visc2) / (2.0_db * visc2 * (rho_r * visc1 + rho_b * visc2))
#else
    cb = fz * hp * hp / (visc1 + visc2)
    cl = fz * hp * (visc1 - visc2) / (2.0_db * visc1 * (visc1 + visc2))
    cr = fz * hp * (visc1 - visc2) / (2.0_db * visc2 * (visc1 + visc2))
#endif
#endif
    ! phi is not reset per node: a solid node keeps the value of the last
    ! fluid node before it, as the reference's loop leaves it, until
    ! solid_phi_average replaces it in runs with obstacles
    do gk = 1, lz
      do gj = 1, ly
        do gi = 1, lx
          press = 0; u = 0; v = 0; w = 0
          if (abs(int(isf(gi, gj, gk))) == 1) then
#if defined(TAYLORGREEN)
            x = 2 * pi_greek * real(gi - 1, db) / real(lx, db)
            y = 2 * pi_greek * real(gj - 1, db) / real(ly, db)
            z = 2 * pi_greek * real(gk - 1, db) / real(lz, db)
            u = uwall * sin(x) * cos(y) * cos(z)
            v = -uwall * cos(x) * sin(y) * cos(z)
            press = ((uwall * uwall / 16.0_db) * (cos(2 * x) + cos(2 * y)) * (2.0_db + cos(2 * z))) / cssq
#elif defined(POISEUILLE)
            dist = real(gi, db) - xc
            if (abs(dist) <= hp) w = fz / (2.0_db * visc1) * (hp * hp - dist * dist)
#elif defined(TWOPOISEUILLE)
            dist = real(gi, db) - xc
            phi = fcut_tanh(dist, 0.0_db, width)
            if (abs(dist) <= hp) then
              wl = -(fz / (2.0_db * visc1)) * dist * dist + cl * dist + cb
              wr = -(fz / (2.0_db * visc2)) * dist * dist + cr * dist + cb
              w = wl * phi + (1.0_db - phi) * wr
            end if
#else
            d(1) = image(real(gi, db) - center(1), lx)
            d(2) = image(real(gj, db) - center(2), ly)
            d(3) = image(real(gk, db) - center(3), lz)
#ifdef LAMBTEST
            dist = sqrt((d(1) * 0.5_db)**2 + d(2)**2 + d(3)**2)
#else
            dist = sqrt(d(1)**2 + d(2)**2 + d(3)**2)
#endif
            phi = fcut_tanh(dist, radius, width)
#ifdef TWOCOMPONENT
#ifdef DENSRATIO
            rhophi = rho_r * phi + (1.0_db - phi) * rho_b
#else
            rhophi = 1.0_db
#endif
            press = phi * (sigma * 2.0_db) / radius / (rhophi * cssq)
#else
            press = 0.01_db * phi
#endif
#endif
          end if
          f(gi, gj, gk, 1) = press
          f(gi, gj, gk, 2) = u; f(gi, gj, gk, 3) = v; f(gi, gj, gk, 4) = w
          f(gi, gj, gk, 5) = u * u; f(gi, gj, gk, 6) = v * v; f(gi, gj, gk, 7) = w * w
          f(gi, gj, gk, 8) = u * v; f(gi, gj, gk, 9) = u * w; f(gi, gj, gk, 10) = v * w
          f(gi, gj, gk, 11) = phi
        end do
      end do
    end do
#if defined(TWOCOMPONENT) && defined(INTERNAL_OBSTACLES)
    call solid_phi_average(f(:, :, :, 11))
#endif
  end subroutine

  ! init_conditions_module.f90:468-542: phi at each solid node (isfluid 0)
  ! becomes the mean of its in-range non-solid lattice neighbours, weighted
  ! 1, 1/sqrt(2), 1/sqrt(3) by link length; a solid node with no such
  ! neighbour gets 0 (the reference divides 0 by 0 there)
  subroutine solid_phi_average(ph)
    real(db), intent(inout) :: ph(:, :, :)
    real(db) :: num, den, s
    integer :: gi, gj, gk, a, b, c, ni, nj, nk, d2
    ! only solid nodes are written and only non-solid ones read: in place
    do gk = 1, lz
      do gj = 1, ly
        do gi = 1, lx
          if (isf(gi, gj, gk) /= 0) cycle
          num = 0; den = 0
          do c = -1, 1
            do b = -1, 1
              do a = -1, 1
                d2 = a * a + b * b + c * c
                if (d2 == 0 .or. (LATTICE == 19 .and. d2 == 3)) cycle
                ni = gi + a; nj = gj + b; nk = gk + c
                if (ni < 1 .or. ni > lx .or. nj < 1 .or. nj > ly .or. nk < 1 .or. nk > lz) cycle
                if (isf(ni, nj, nk) == 0) cycle
                s = 1.0_db
                if (d2 == 2) s = 1.0_db / sqrt(2.0_db)
                if (d2 == 3) s = 1.0_db / sqrt(3.0_db)
                num = num + s * ph(ni, nj, nk); den = den + s
              end do
            end do
          end do
          ph(gi, gj, gk) = 0
          if (den > 0) ph(gi, gj, gk) = num / den
        end do
      end do
    end do
  end subroutine

  subroutine download(id, a)
    integer(c_int), intent(in) :: id
    real(db), intent(out), target :: a(:, :, :)
    integer(c_int) :: code
    code = lbf_download_field(ctx, id, c_loc(a(1, 1, 1)), LBF_DTYPE_F64)
    call check(code, 'lbf_download_field')
  end subroutine

  real(db) function kinetic_energy()
    real(db) :: s
    integer(c_int) :: id
    s = 0
    do id = LBF_FIELD_U, LBF_FIELD_W
      call download(id, buf)
      s = s + sum(buf**2)
    end do
    kinetic_energy = 0.5_db * s / (real(lx, db) * real(ly, db) * real(lz, db))
  end function

  ! statistics.f90 diagnostics
  subroutine report(st)
    integer, intent(in) :: st
    real(db) :: e, k2, eth, rad, left, right, wsum, wl, wr, p_in, p_out, rho_in, rho_out, ph_in, ph_out
    integer :: ic, jc, kc, ii
    integer(c_int) :: code
    code = lbf_reduce_fp64(ctx, red)
    call check(code, 'lbf_reduce_fp64')
#if defined(TAYLORGREEN)
    e = kinetic_energy()
    k2 = (2 * pi_greek / real(lx, db))**2
    eth = exp(-6.0_db * visc1 * k2 * real(st, db))
    write (6, '(a,i10,a,es20.12,a,es20.12)') 'TAYLORGREEN step', st, ' E/E0', e / e0, ' theory', eth
#elif defined(TWOCOMPONENT)
    ! statistics.f90:184-275: interface positions from phi(1-phi) weights
    ! along x through (ly/2, lz/2), pressure inside at the centre node and
    ! outside at the probe node (iprobe, jprobe, kprobe)
    call download(LBF_FIELD_PHI, buf)
    if (.not. allocated(buf2)) allocate (buf2(lx, ly, lz))
    call download(LBF_FIELD_P, buf2)
    jc = max(1, ly / 2); kc = max(1, lz / 2)
    left = 0; right = 0; wl = 0; wr = 0
    do ii = 1, lx
      wsum = buf(ii, jc, kc) * (1.0_db - buf(ii, jc, kc))
      if (ii <= lx / 2) then
        left = left + ii * wsum; wl = wl + wsum
      else
        right = right + ii * wsum; wr = wr + wsum
      end if
    end do
    rad = 0
    if (wl > 0 .and. wr > 0) rad = right / wr - 0.5_db * (right / wr + left / wl)
    ic = max(1, min(lx, nint(center(1)))); jc = max(1, min(ly, nint(center(2)))); kc = max(1, min(lz, nint(center(3))))
    ph_in = buf(ic, jc, kc)
    rho_in = rho_r * ph_in + (1.0_db - ph_in) * rho_b
    p_in = rho_in * cssq * buf2(ic, jc, kc)
    ic = max(1, min(lx, iprobe)); jc = max(1, min(ly, jprobe)); kc = max(1, min(lz, kprobe))
    ph_out = buf(ic, jc, kc)
    rho_out = rho_r * ph_out + (1.0_db - ph_out) * rho_b
    p_out = rho_out * cssq * buf2(ic, jc, kc)
    write (6, '(a,i10,a,es20.12,a,es20.12,a,es20.12,a,es20.12)') 'LAPLACE step', st, ' sigma_eff', &
      (p_in - p_out) * rad * 0.5_db, ' delta_p', p_in - p_out, ' radius', rad, ' phase_mass', red(5)
#if defined(LAMBTEST)
    call lamb_line(st)
#endif
#else
    write (6, '(a,i10,a,es20.12)') 'MASS step', st, ' sum_P', red(1)
#endif
  end subroutine

  ! RAW files: stream, unformatted, real(4), Fortran order (print_module.f90:826-1007)
  subroutine read_raw(name, a)
    character(len=*), intent(in) :: name
    real(db), intent(out) :: a(:, :, :)
    real(4), allocatable :: a4(:, :, :)
    logical :: ex
    inquire (file=name, exist=ex)
    if (.not. ex) then
      write (6, '(a,a)') 'lbfast_host: unable to find the file: ', name
      stop 2
    end if
    allocate (a4(size(a, 1), size(a, 2), size(a, 3)))
    open (unit=345, file=name, status='old', action='read', access='stream', form='unformatted')
    read (345) a4
    close (345)
    a = real(a4, db)
  end subroutine

  ! every stepskip-th node along each axis (copy_print, print_module.f90:638-776)
  subroutine skip_copy(a, o)
    real(db), intent(in) :: a(:, :, :)
    real(4), intent(out) :: o(:, :, :)
    integer :: i, j, k
    do k = 1, size(o, 3)
      do j = 1, size(o, 2)
        do i = 1, size(o, 1)
          o(i, j, k) = real(a(i * stepskip, j * stepskip, k * stepskip), 4)
        end do
      end do
    end do
  end subroutine

  character(len=6) function fmtnumb(n)
    integer, intent(in) :: n
    write (fmtnumb, '(i6.6)') max(0, min(999999, n))
  end function

  ! copy_print (print_module.f90:638-776): the print observables at every
  ! stepskip-th node: rho is rho_r phi + (1 - phi) rho_b with DENSRATIO (phi
  ! with PRINTPHI), phi in a two-component build without it, P in a
  ! one-component build; vel holds u, v, w interleaved per node; press is P
  subroutine copy_print(rp, vp, pp)
    real(4), intent(out) :: rp(:, :, :), vp(:, :, :, :), pp(:, :, :)
    integer :: c
#if defined(TWOCOMPONENT)
    call download(LBF_FIELD_PHI, buf)
#if defined(DENSRATIO) && !defined(PRINTPHI)
    buf = rho_r * buf + (1.0_db - buf) * rho_b
#endif
#else
    call download(LBF_FIELD_P, buf)
#endif
    call skip_copy(buf, rp)
    do c = 1, 3
      call download(int(LBF_FIELD_U + c - 1, c_int), buf)
      call skip_copy(buf, pp)
      vp(c, :, :, :) = pp
    end do
    call download(LBF_FIELD_P, buf)
    call skip_copy(buf, pp)
  end subroutine

  ! print_raw_sync (print_module.f90:2007-2180) and print_vtk_sync: frames
  ! output/out_rho_NNNNNN, out_vel_NNNNNN and, with WRITEPRESS in a two-
  ! component build, out_press_NNNNNN; real(4), stream access.  lraw writes
  ! .raw files (with an .xdmf description beside each under DOXDMF), lvtk
  ! VTK image data (.vti, raw appended data)
  subroutine write_frame(fr)
    integer, intent(in) :: fr
    real(4), allocatable :: rp(:, :, :), vp(:, :, :, :), pp(:, :, :)
    integer :: nxs, nys, nzs
    nxs = lx / stepskip; nys = ly / stepskip; nzs = lz / stepskip
    allocate (rp(nxs, nys, nzs), vp(3, nxs, nys, nzs), pp(nxs, nys, nzs))
    call copy_print(rp, vp, pp)
    if (lraw) then
      call write_raw4('output/out_rho_'//fmtnumb(fr)//'.raw', rp)
      call write_raw4v('output/out_vel_'//fmtnumb(fr)//'.raw', vp)
#if defined(DOXDMF)
      call write_xdmf('output/out_rho_'//fmtnumb(fr), 'rho', [nxs, nys, nzs], 1)
      call write_xdmf('output/out_vel_'//fmtnumb(fr), 'vel', [nxs, nys, nzs], 3)
#endif
#if defined(TWOCOMPONENT) && defined(WRITEPRESS)
      call write_raw4('output/out_press_'//fmtnumb(fr)//'.raw', pp)
#if defined(DOXDMF)
      call write_xdmf('output/out_press_'//fmtnumb(fr), 'press', [nxs, nys, nzs], 1)
#endif
#endif
    end if
    if (lvtk) then
      call write_vti('output/out_rho_'//fmtnumb(fr)//'.vti', 'rho', rp, [nxs, nys, nzs], 1)
      call write_vti('output/out_vel_'//fmtnumb(fr)//'.vti', 'vel', vp, [nxs, nys, nzs], 3)
#if defined(TWOCOMPONENT) && defined(WRITEPRESS)
      call write_vti('output/out_press_'//fmtnumb(fr)//'.vti', 'press', pp, [nxs, nys, nzs], 1)
#endif
    end if
  end subroutine

  ! driver_print_raw_sync2D / print_raw_sync2D (print_module.f90:2307-2660):
  ! for each plane l, the slice normal to axis ndir(l) at node npoint(l) of
  ! the print observables, as output/out_<ndir>_<npoint>_rho_NNNNNN.raw,
  ! ..._vel_... (and ..._press_... with WRITEPRESS), with .xdmf sidecars under
  ! DOXDMF
  subroutine write_frame_2d(fr)
    integer, intent(in) :: fr
    real(4), allocatable :: rp(:, :, :), vp(:, :, :, :), pp(:, :, :), s1(:, :, :), s3(:, :, :, :), s2(:, :, :)
    integer :: nxs, nys, nzs, l, pt, d(3)
    character(len=64) :: stem
    nxs = lx / stepskip; nys = ly / stepskip; nzs = lz / stepskip
    allocate (rp(nxs, nys, nzs), vp(3, nxs, nys, nzs), pp(nxs, nys, nzs))
    call copy_print(rp, vp, pp)
    do l = 1, nplanes
      pt = npoint(l) / stepskip
      select case (ndir(l))
      case (1)
        d = [1, nys, nzs]
        if (pt < 1 .or. pt > nxs) call plane_error(l)
        s1 = rp(pt:pt, :, :); s2 = pp(pt:pt, :, :); s3 = vp(:, pt:pt, :, :)
      case (2)
        d = [nxs, 1, nzs]
        if (pt < 1 .or. pt > nys) call plane_error(l)
        s1 = rp(:, pt:pt, :); s2 = pp(:, pt:pt, :); s3 = vp(:, :, pt:pt, :)
      case (3)
        d = [nxs, nys, 1]
        if (pt < 1 .or. pt > nzs) call plane_error(l)
        s1 = rp(:, :, pt:pt); s2 = pp(:, :, pt:pt); s3 = vp(:, :, :, pt:pt)
      case default
        write (6, '(a)') 'wrong argument in ndir for 2d plane print: only 1 to 3 is permitted'
        stop 2
      end select
      write (stem, '(a,i0,a,i0,a)') 'output/out_', ndir(l), '_', npoint(l), '_'
      call write_raw4(trim(stem)//'rho_'//fmtnumb(fr)//'.raw', s1)
      call write_raw4v(trim(stem)//'vel_'//fmtnumb(fr)//'.raw', s3)
#if defined(DOXDMF)
      call write_xdmf(trim(stem)//'rho_'//fmtnumb(fr), 'rho', d, 1)
      call write_xdmf(trim(stem)//'vel_'//fmtnumb(fr), 'vel', d, 3)
#endif
#if defined(TWOCOMPONENT) && defined(WRITEPRESS)
      call write_raw4(trim(stem)//'press_'//fmtnumb(fr)//'.raw', s2)
#if defined(DOXDMF)
      call write_xdmf(trim(stem)//'press_'//fmtnumb(fr), 'press', d, 1)
#endif
#endif
    end do
  end subroutine

  subroutine plane_error(l)
    integer, intent(in) :: l
    write (6, '(a,i0,a,i0,a,i0)') 'lbfast_host: 2-D plane ', l, ' at node ', npoint(l), &
      ' lies outside the domain along axis ', ndir(l)
    stop 2
  end subroutine

  ! XDMF description of a RAW file (print_module.f90:2020-2064): a uniform
  ! 3DCoRectMesh, origin 0, spacing 1, dimensions slowest axis first, one
  ! node-centred attribute (a vector when ncomp = 3) in little-endian float32
  subroutine write_xdmf(stem, name, d, ncomp)
    character(len=*), intent(in) :: stem, name
    integer, intent(in) :: d(3), ncomp
    character(len=64) :: dims
    integer :: u, k
    write (dims, '(i0,a,i0,a,i0)') d(3), ' ', d(2), ' ', d(1)
    k = index(stem, '/', back=.true.)
    open (newunit=u, file=stem//'.xdmf', status='replace', action='write')
    write (u, '(a)') '<?xml version="1.0" ?>'
    write (u, '(a)') '<Xdmf Version="3.0">'
    write (u, '(a)') '  <Domain>'
    write (u, '(a)') '    <!-- Uniform Grid -->'
    if (ncomp == 3) then
      write (u, '(a)') '    <Grid Name="VelocityField" GridType="Uniform">'
    else
      write (u, '(a)') '    <Grid Name="ScalarField" GridType="Uniform">'
    end if
    write (u, '(a)') ''
    write (u, '(a)') '      <!-- Topology: 3D grid -->'
    write (u, '(3a)') '      <Topology TopologyType="3DCoRectMesh" Dimensions="', trim(dims), '"/>'
    write (u, '(a)') ''
    write (u, '(a)') '      <!-- Geometry: origin and spacing (dx, dy, dz) -->'
    write (u, '(a)') '      <Geometry GeometryType="ORIGIN_DXDYDZ">'
    write (u, '(a)') '        <DataItem Dimensions="3" Format="XML" Precision="4">0.0 0.0 0.0</DataItem>'
    write (u, '(a)') '        <DataItem Dimensions="3" Format="XML" Precision="4">1.0 1.0 1.0</DataItem>'
    write (u, '(a)') '      </Geometry>'
    write (u, '(a)') ''
    if (ncomp == 3) then
      write (u, '(a)') '      <!-- Vector field associated with the grid -->'
      write (u, '(3a)') '      <Attribute Name="', name, '" AttributeType="Vector" Center="Node">'
      write (u, '(3a)') '        <DataItem Format="Binary" NumberType="Float" Precision="4" Dimensions="', &
        trim(dims)//' 3', '" Endian="Little">'
    else
      write (u, '(a)') '      <!-- Scalar field associated with the grid -->'
      write (u, '(3a)') '      <Attribute Name="', name, '" AttributeType="Scalar" Center="Node">'
      write (u, '(3a)') '        <DataItem Format="Binary" NumberType="Float" Precision="4" Dimensions="', &
        trim(dims), '" Endian="Little">'
    end if
    write (u, '(2a)') '          ', stem(k + 1:)//'.raw'
    write (u, '(a)') '        </DataItem>'
    write (u, '(a)') '      </Attribute>'
    write (u, '(a)') ''
    write (u, '(a)') '    </Grid>'
    write (u, '(a)') '  </Domain>'
    write (u, '(a)') '</Xdmf>'
    close (u)
  end subroutine

  ! VTK XML image data with raw appended float32 data (header_vtk /
  ! footer_vtk / print_vtk_sync, print_module.f90:22-244): extent 1..n per
  ! axis, one point-data array of ncomp components, the appended block "_",
  ! its byte count as int32, the values (components interleaved per node)
  subroutine write_vti(name, var, a, d, ncomp)
    character(len=*), intent(in) :: name, var
    real(4), intent(in) :: a(*)
    integer, intent(in) :: d(3), ncomp
    character(len=1), parameter :: nl = char(10)
    character(len=96) :: ext
    integer :: u
    integer(4) :: nbytes
    write (ext, '(i0,a,i0,a,i0,a,i0,a,i0,a,i0)') 1, ' ', d(1), ' ', 1, ' ', d(2), ' ', 1, ' ', d(3)
    nbytes = int(4, 4) * int(ncomp, 4) * int(d(1), 4) * int(d(2), 4) * int(d(3), 4)
    open (newunit=u, file=name, status='replace', action='write', access='stream', form='unformatted')
    write (u) '<?xml version="1.0"?>'//nl
    write (u) '<VTKFile type="ImageData" version="0.1" byte_order="LittleEndian">'//nl
    write (u) '  <ImageData WholeExtent="'//trim(ext)//'">'//nl
    write (u) '    <Piece Extent="'//trim(ext)//'">'//nl
    write (u) '      <PointData>'//nl
    write (u) '        <DataArray type="Float32" Name="'//var//'" NumberOfComponents="'// &
      achar(48 + ncomp)//'" format="appended" offset="0"/>'//nl
    write (u) '      </PointData>'//nl
    write (u) '    </Piece>'//nl
    write (u) '  </ImageData>'//nl
    write (u) '  <AppendedData encoding="raw">'//nl
    write (u) '_'
    write (u) nbytes
    write (u) a(1:ncomp * d(1) * d(2) * d(3))
    write (u) nl//'  </AppendedData>'//nl//'</VTKFile>'//nl
    close (u)
  end subroutine

  ! print_raw_isfluid (print_module.f90:1907-1990): the classified isfluid as
  ! real(4), written once (kept when the file already exists)
  subroutine write_isfluid_frame()
    real(4), allocatable :: o(:, :, :)
    logical :: ex
    inquire (file='output/out_isfluid_'//fmtnumb(iframe)//'.raw', exist=ex)
    if (ex) return
    buf = real(isf, db)
    allocate (o(lx / stepskip, ly / stepskip, lz / stepskip))
    call skip_copy(buf, o)
    call write_raw4('output/out_isfluid_'//fmtnumb(iframe)//'.raw', o)
  end subroutine

  subroutine write_raw4(name, o)
    character(len=*), intent(in) :: name
    real(4), intent(in) :: o(:, :, :)
    open (unit=345, file=name, status='replace', action='write', access='stream', form='unformatted')
    write (345) o
    close (345)
  end subroutine

  subroutine write_raw4v(name, o)
    character(len=*), intent(in) :: name
    real(4), intent(in) :: o(:, :, :, :)
    open (unit=345, file=name, status='replace', action='write', access='stream', form='unformatted')
    write (345) o
    close (345)
  end subroutine

#if defined(LAMBTEST) && defined(TWOCOMPONENT)
  ! open_lamb / print_lamb (statistics.f90:277-471): interface positions of
  ! the print observable r (phi with PRINTPHI, rho_r phi + (1 - phi) rho_b
  ! otherwise) as r (1 - r)-weighted mean coordinates of each half of the line
  ! along z through (lx/2, ly/2) and along x through (ly/2, lz/2); one line of
  ! lamb.dat per call: step, (Lz - Lx)/(Lz + Lx), Lx, Lz, the two centres,
  ! and the z half-axis z_int - z_cm
  subroutine lamb_line(st)
    integer, intent(in) :: st
    real(db) :: zl, zln, zr, zrn, xl, xln, xr, xrn, w, r, lamb_x, lamb_z, cmx, cmz
    integer :: gi, gj, gk
    call download(LBF_FIELD_PHI, buf)
#if defined(DENSRATIO) && !defined(PRINTPHI)
    buf = rho_r * buf + (1.0_db - buf) * rho_b
#endif
    gi = max(1, lx / 2); gj = max(1, ly / 2)
    zl = 0; zln = 0; zr = 0; zrn = 0
    do gk = 1, lz
      r = buf(gi, gj, gk); w = r * (1.0_db - r)
      if (gk <= lz / 2) then
        zl = zl + real(gk, db) * w; zln = zln + w
      else
        zr = zr + real(gk, db) * w; zrn = zrn + w
      end if
    end do
    gj = max(1, ly / 2); gk = max(1, lz / 2)
    xl = 0; xln = 0; xr = 0; xrn = 0
    do gi = 1, lx
      r = buf(gi, gj, gk); w = r * (1.0_db - r)
      if (gi <= lx / 2) then
        xl = xl + real(gi, db) * w; xln = xln + w
      else
        xr = xr + real(gi, db) * w; xrn = xrn + w
      end if
    end do
    zl = zl / zln; zr = zr / zrn; xl = xl / xln; xr = xr / xrn
    lamb_z = zr - zl; cmz = 0.5_db * (zr + zl)
    lamb_x = xr - xl; cmx = 0.5_db * (xr + xl)
    write (142, '(i8,7g16.8)') st, (lamb_z - lamb_x) / (lamb_z + lamb_x), lamb_x, lamb_z, cmx, cmz, zr - cmz
    flush (142)
  end subroutine

  ! compare_benchmark, Lamb (statistics.f90:959-1114): the peaks of the z
  ! half-axis in lamb.dat (local maxima above the radius, at least 2000 steps
  ! apart), their mean spacing T_num and omega_num = 2 pi / T_num, against
  ! Lamb's l = 2 frequency sqrt(24 sigma / (R^3 (2 rho_b + 3 rho_r))) with the
  ! viscous correction of the two fluids; lamb_theory.dat holds R + a0 cos
  ! of the theory at the recorded steps
  subroutine lamb_period()
    integer, parameter :: min_peak_dist = 2000
    integer, allocatable :: mystep(:), pstep(:)
    real(db), allocatable :: dosc(:), pval(:)
    real(db) :: r_eq, mu1, mu2, om, chi, om_c, period_th, tmean, om_num, amp0, t0, a, b, c, d, e
    integer :: maxn, myn, ios, npeaks, k, u
    r_eq = radius
    mu1 = rho_r * visc1
    mu2 = rho_b * visc2
    om = sqrt(24.0_db * sigma / (r_eq**3 * (2.0_db * rho_b + 3.0_db * rho_r)))
    chi = ((2.0_db * 2.0_db + 1.0_db)**2 * sqrt(mu1 * mu2 * rho_r * rho_b)) / &
          (sqrt(2.0_db) * r_eq * (2.0_db * rho_b + 3.0_db * rho_r) * (sqrt(mu1 * rho_r) + sqrt(mu2 * rho_b)))
    om_c = om - 0.5_db * chi * sqrt(om) + 0.25_db * chi**2
    period_th = 2.0_db * pi_greek / om_c
    maxn = ntot / max(1, stamp_term) + 1
    allocate (mystep(maxn), dosc(maxn), pstep(maxn), pval(maxn))
    open (newunit=u, file='lamb.dat', status='old', action='read', iostat=ios)
    if (ios /= 0) then
      write (6, '(a)') 'Error in opening lamb.dat'
      return
    end if
    myn = 0
    do
      if (myn >= maxn) exit
      read (u, '(i8,7g16.8)', iostat=ios) k, a, b, c, d, e, amp0
      if (ios /= 0) exit
      myn = myn + 1
      mystep(myn) = k
      dosc(myn) = amp0
    end do
    close (u)
    if (myn < 3) then
      write (6, '(a)') 'Too few points in lamb.dat'
      return
    end if
    amp0 = dosc(1) - r_eq
    t0 = real(mystep(1), db)
    npeaks = 0
    do k = 2, myn - 1
      if (dosc(k) > dosc(k - 1) .and. dosc(k) >= dosc(k + 1) .and. dosc(k) > r_eq) then
        if (npeaks == 0) then
          npeaks = 1; pstep(1) = mystep(k); pval(1) = dosc(k)
        else if (mystep(k) - pstep(npeaks) >= min_peak_dist) then
          npeaks = npeaks + 1; pstep(npeaks) = mystep(k); pval(npeaks) = dosc(k)
        end if
      end if
    end do
    open (newunit=u, file='lamb_theory.dat', status='replace', action='write')
    do k = 1, myn
      write (u, '(i8,g16.8)') mystep(k), r_eq + amp0 * cos(om_c * (real(mystep(k), db) - t0))
    end do
    close (u)
    if (npeaks < 2) then
      write (6, '(a)') 'Not enough peaks found to estimate omega.'
      return
    end if
    tmean = real(pstep(npeaks) - pstep(1), db) / real(npeaks - 1, db)
    om_num = 2.0_db * pi_greek / tmean
    write (6, '(a,i10)') 'Expected number of points       = ', maxn
    write (6, '(a,i10)') 'Number of points read           = ', myn
    write (6, '(a,i10)') 'Number of peaks found           = ', npeaks
    write (6, '(a,f20.10)') 'Equivalent radius               = ', r_eq
    write (6, '(a,f20.10)') 'Theoretical period              = ', period_th
    write (6, '(a,f20.10)') 'Theoretical omega               = ', om_c
    write (6, '(a,f20.10)') 'Mean numerical period T_num     = ', tmean
    write (6, '(a,f20.10)') 'Angular frequency omega_num     = ', om_num
    write (6, '(a,f20.10)') 'zint0-zcm                       = ', dosc(1)
    write (6, '(a,f20.10)') 'amp0                            = ', amp0
    do k = 1, npeaks
      write (6, '(a,i6,a,i10,a,g16.8)') 'peak ', k, ' step=', pstep(k), ' zint-zcm=', pval(k)
    end do
    write (6, '(a,2es20.12)') 'LAMB omega_num omega_th ', om_num, om_c
  end subroutine
#endif

#if defined(POISEUILLE) || (defined(TWOPOISEUILLE) && defined(TWOCOMPONENT))
  ! compare_benchmark, Poiseuille (statistics.f90:688-916): w averaged over
  ! each x column (zero on solid nodes) against the analytic channel profile
  ! -- the parabola fz/(2 nu) (H^2 - d^2), or the two-fluid profile of the
  ! layers split by fcut_tanh at the centre -- with H = (lx-2)/2 under
  ! BOUNCE_BACK, (lx-1)/2 otherwise; written to plot_poiseuille.dat or
  ! plot_twopoiseuille.dat (x, numerical, analytic), plus a summary line
  subroutine poiseuille_profile()
    real(db), allocatable :: w_num(:), w_pois(:), i_num(:)
    real(db) :: hp, xc, dist, cb, cl, cr, wl, wr, ph, en, ed
    integer :: gi, gj, gk, u
    character(len=32) :: fname
    allocate (w_num(lx), w_pois(lx), i_num(lx))
#ifdef BOUNCE_BACK
    hp = 0.5_db * real(lx - 2, db)
#else
    hp = 0.5_db * real(lx - 1, db)
#endif
    xc = 0.5_db * real(lx + 1, db)
#if defined(TWOPOISEUILLE)
#ifdef DENSRATIO
    cb = fz * hp * hp * (rho_r + rho_b) / (2.0_db * (rho_r * visc1 + rho_b * visc2))
    cl = fz * hp * rho_b * (visc1 - visc2) / (2.0_db * visc1 * (rho_r * visc1 + rho_b * visc2))
    cr = fz * hp * rho_r * (visc1 - visc2) / (2.0_db * visc2 * (rho_r * visc1 + rho_b * visc2))
#else
    cb = fz * hp * hp / (visc1 + visc2)
    cl = fz * hp * (visc1 - visc2) / (2.0_db * visc1 * (visc1 + visc2))
    cr = fz * hp * (visc1 - visc2) / (2.0_db * visc2 * (visc1 + visc2))
#endif
    fname = 'plot_twopoiseuille.dat'
#else
    fname = 'plot_poiseuille.dat'
#endif
    do gi = 1, lx
      dist = real(gi, db) - xc
      w_pois(gi) = 0
      if (abs(dist) <= hp) then
#if defined(TWOPOISEUILLE)
        ph = fcut_tanh(dist, 0.0_db, width)
        wl = -(fz / (2.0_db * visc1)) * dist * dist + cl * dist + cb
        wr = -(fz / (2.0_db * visc2)) * dist * dist + cr * dist + cb
        w_pois(gi) = wl * ph + (1.0_db - ph) * wr
#else
        w_pois(gi) = fz / (2.0_db * visc1) * (hp * hp - dist * dist)
#endif
      end if
    end do
    call download(LBF_FIELD_W, buf)
    w_num = 0; i_num = 0
    do gk = 1, lz
      do gj = 1, ly
        do gi = 1, lx
          if (isf(gi, gj, gk) /= 0) w_num(gi) = w_num(gi) + buf(gi, gj, gk)
          i_num(gi) = i_num(gi) + 1.0_db
        end do
      end do
    end do
    w_num = w_num / i_num
    open (newunit=u, file=trim(fname), status='replace', action='write')
#ifdef BOUNCE_BACK
    write (u, '(3g16.8)') 1.5_db, w_num(1), w_pois(1)
    do gi = 2, lx - 1
      write (u, '(3g16.8)') real(gi, db), w_num(gi), w_pois(gi)
    end do
    write (u, '(3g16.8)') real(lx, db) - 0.5_db, w_num(lx), w_pois(lx)
#else
    do gi = 1, lx
      write (u, '(3g16.8)') real(gi, db), w_num(gi), w_pois(gi)
    end do
#endif
    close (u)
    en = sqrt(sum((w_num - w_pois)**2)); ed = sqrt(sum(w_pois**2))
    write (6, '(3a,es20.12,a,es20.12,a,es20.12)') 'POISEUILLE ', trim(fname), ' rel_l2', en / max(ed, tiny(1.0_db)), &
      ' max_num', maxval(w_num), ' max_analytic', maxval(w_pois)
  end subroutine
#endif

  ! output/restart.raw (print_module.f90:1474-1893): the ten moments in the
  ! order 1,2,3,4,5,8,9,6,10,7 (P, u, v, w, Pxx, Pxy, Pxz, Pyy, Pyz, Pzz),
  ! then phi in a two-component build, as real(4) arrays, then iframe and
  ! iframe2D
  subroutine write_restart(fr, fr2)
    integer, intent(in) :: fr, fr2
    integer, parameter :: order(10) = [1, 2, 3, 4, 5, 8, 9, 6, 10, 7]
    integer :: n
    open (unit=345, file='output/restart.raw', status='replace', action='write', access='stream', &
          form='unformatted')
    do n = 1, 10
      call download(int(order(n), c_int), buf)
      write (345) real(buf, 4)
    end do
    if (two) then
      call download(LBF_FIELD_PHI, buf)
      write (345) real(buf, 4)
    end if
    write (345) fr, fr2
    close (345)
  end subroutine

  subroutine read_restart(f, fr, fr2)
    real(db), intent(inout) :: f(:, :, :, :)
    integer, intent(out) :: fr, fr2
    integer, parameter :: order(10) = [1, 2, 3, 4, 5, 8, 9, 6, 10, 7]
    real(4), allocatable :: a4(:, :, :)
    integer :: n
    logical :: ex
    inquire (file='output/restart.raw', exist=ex)
    if (.not. ex) then
      write (6, '(a)') 'lbfast_host: unable to find the file: output/restart.raw'
      stop 2
    end if
    allocate (a4(lx, ly, lz))
    open (unit=345, file='output/restart.raw', status='old', action='read', access='stream', form='unformatted')
    do n = 1, 10
      read (345) a4
      f(:, :, :, order(n)) = real(a4, db)
    end do
    if (two) then
      read (345) a4
      f(:, :, :, 11) = real(a4, db)
    end if
    read (345) fr, fr2
    close (345)
  end subroutine

end program lbfast_host
