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
!    sigma_eff = (p_in - p_out) R / 2, global mass;
!  * performance as LBFAST.f90:500 (MLUPS), RAW output (lraw) and the restart
!    file (lwriterestart / lrestart) in the reference's formats.
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
  p%bounce_back = 0; p%smagorinsky = 0; p%repulsive_flux = 0
#ifdef BOUNCE_BACK
  p%bounce_back = 1
#endif
#ifdef SMAGORINSKI
  p%smagorinsky = 1
#endif
#ifdef REPULSIVE_FLUX
  p%repulsive_flux = 1
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
  rc = lbf_params_derive(p)
  call check(rc, 'lbf_params_derive')
  two = p%twocomponent /= 0
  write (6, '(a,3i6,a,i3,a,i2,a,i2,a,i2)') 'lbfast_host: lx ly lz', lx, ly, lz, '  lattice', p%lattice, &
    '  highorder', p%highorder, '  twocomponent', p%twocomponent, '  precision', p%precision
  write (6, '(a,7i2,i4)') 'lbfast_host: switches densratio incomp obstacles buoyancy bounce_back smagorinsky'// &
    ' repulsive pbc', p%densratio, p%interface_incomp, p%internal_obstacles, p%buoyancy_forcing, &
    p%bounce_back, p%smagorinsky, p%repulsive_flux, p%pbc(1) * 100 + p%pbc(2) * 10 + p%pbc(3)
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
  if (lprint .and. lraw .and. .not. lrestart) call write_frame(iframe)

  ! ---- time loop (integrator_module.f90:266-827) ------------------------------------------
  ! The reference advances FLIP/FLOP pairs: 2*ceil(nsteps/2) steps.  Steps
  ! are enqueued in chunks that end on the next step at which output,
  ! statistics or the wall-clock check is due.
  e0 = kinetic_energy()
  call report(0)
  ntot = 2 * ((max(nsteps, 0) + 1) / 2)
  stop_early = .false.
  call system_clock(c0, crate)
  step = 0
  do while (step < ntot)
    nxt = ntot
    if (stamp > 0 .and. lprint) nxt = min(nxt, (step / stamp + 1) * stamp)
    if (stamp_term > 0) nxt = min(nxt, (step / stamp_term + 1) * stamp_term)
    if (time_limit > 0 .and. every_time_check > 0) nxt = min(nxt, (step / every_time_check + 1) * every_time_check)
    chunk = nxt - step
    rc = lbf_step(ctx, int(chunk, c_int32_t))
    call check(rc, 'lbf_step')
    step = nxt
    if (lprint .and. stamp > 0) then
      if (mod(step, stamp) == 0) then
        iframe = iframe + 1
        if (lraw) call write_frame(iframe)
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
    ! fluid node before it, as the reference's loop leaves it
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

  ! print_raw_sync (print_module.f90:2007-2180): output/out_rho_NNNNNN.raw,
  ! out_vel_NNNNNN.raw (u, v, w interleaved per node) and, with WRITEPRESS
  ! in a two-component build, out_press_NNNNNN.raw; real(4), stream access.
  ! rho is rho_r phi + (1 - phi) rho_b with DENSRATIO (phi with PRINTPHI),
  ! phi in a two-component build without it, P in a one-component build.
  subroutine write_frame(fr)
    integer, intent(in) :: fr
    real(4), allocatable :: o(:, :, :), vel(:, :, :, :)
    integer :: nxs, nys, nzs, c
    nxs = lx / stepskip; nys = ly / stepskip; nzs = lz / stepskip
    allocate (o(nxs, nys, nzs), vel(3, nxs, nys, nzs))
#if defined(TWOCOMPONENT)
    call download(LBF_FIELD_PHI, buf)
#if defined(DENSRATIO) && !defined(PRINTPHI)
    buf = rho_r * buf + (1.0_db - buf) * rho_b
#endif
#else
    call download(LBF_FIELD_P, buf)
#endif
    call skip_copy(buf, o)
    call write_raw4('output/out_rho_'//fmtnumb(fr)//'.raw', o)
    do c = 1, 3
      call download(int(LBF_FIELD_U + c - 1, c_int), buf)
      call skip_copy(buf, o)
      vel(c, :, :, :) = o
    end do
    open (unit=345, file='output/out_vel_'//fmtnumb(fr)//'.raw', status='replace', action='write', &
          access='stream', form='unformatted')
    write (345) vel
    close (345)
#if defined(TWOCOMPONENT) && defined(WRITEPRESS)
    call download(LBF_FIELD_P, buf)
    call skip_copy(buf, o)
    call write_raw4('output/out_press_'//fmtnumb(fr)//'.raw', o)
#endif
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
