! lbfast_hip.f90 -- iso_c_binding interface to include/lbfast_hip.h.
!
! The C-ABI replaces the reference's CUDA driver module
! (lb_cuda_driver_module.f90); the mapping of its subroutines to these
! functions is listed in include/lbfast_hip.h.  lbf_params mirrors the C
! struct member by member (18 int32, 33 real(8), then the open-boundary
! block: 8 int32 and 24 real(8); 544 bytes).
module lbfast_hip
  use iso_c_binding
  implicit none

  integer(c_int), parameter :: LBF_OK = 0
  integer(c_int), parameter :: LBF_PREC_FP64 = 0, LBF_PREC_FP32 = 1, LBF_PREC_MIXED = 2, &
                               LBF_PREC_FP16S = 3, LBF_PREC_BF16S = 4
  integer(c_int), parameter :: LBF_STAGE_NORM = 1, LBF_STAGE_PHI_INT_BOUNDARY = 2, &
                               LBF_STAGE_DIV_THETA_N = 3, LBF_STAGE_MOMENTS = 4, LBF_STAGE_FUSED = 5, &
                               LBF_STAGE_UPDATE_PHI = 6, LBF_STAGE_LB_INT_BOUNDARY = 7, &
                               LBF_STAGE_THINFILM_SCAN = 8, LBF_STAGE_REPULSIVE_FLUX = 9, &
                               LBF_STAGE_PHI_LAGRANGE = 10, LBF_STAGE_OPEN_BC_PHI = 11, LBF_STAGE_OPEN_BC_H = 12
  integer(c_int), parameter :: LBF_FIELD_P = 1, LBF_FIELD_U = 2, LBF_FIELD_V = 3, LBF_FIELD_W = 4, &
                               LBF_FIELD_PXX = 5, LBF_FIELD_PHI = 11, LBF_FIELD_RHOPHI = 40
  integer(c_int), parameter :: LBF_DTYPE_F64 = 0, LBF_DTYPE_F32 = 1

  type, bind(C) :: lbf_params
    integer(c_int32_t) :: nx, ny, nz, lattice, highorder, twocomponent, densratio, interface_incomp
    integer(c_int32_t) :: internal_obstacles, buoyancy_forcing, pbc(3), precision
    integer(c_int32_t) :: bounce_back, smagorinsky, repulsive_flux, bcphiflux
    real(c_double) :: fx, fy, fz, visc1, visc2, rho_r, rho_b, tau_diff, sigma, width, wettab_r
    real(c_double) :: tau1, omega, tau2, beta, kapp, sharp_c, store_scale(10)
    real(c_double) :: a_rep, q_th, cos_opp_t, reserved_d
    ! open boundaries; faces x(1) x(2) y(1) y(2) z(1) z(2)
    integer(c_int32_t) :: openbc, openbc_type(6), reserved_i2
    real(c_double) :: openbc_press(6), openbc_u(6), openbc_v(6), openbc_w(6)
  end type

  interface
    subroutine lbf_params_default(p) bind(C, name="lbf_params_default")
      import
      type(lbf_params), intent(inout) :: p
    end subroutine
    integer(c_int) function lbf_params_derive(p) bind(C, name="lbf_params_derive")
      import
      type(lbf_params), intent(inout) :: p
    end function
    integer(c_int) function lbf_classify_isfluid(p, inp, outp) bind(C, name="lbf_classify_isfluid")
      import
      type(lbf_params), intent(in) :: p
      integer(c_int8_t), intent(in) :: inp(*)
      integer(c_int8_t), intent(out) :: outp(*)
    end function
    integer(c_int) function lbf_create(p, ctx) bind(C, name="lbf_create")
      import
      type(lbf_params), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function
    subroutine lbf_destroy(ctx) bind(C, name="lbf_destroy")
      import
      type(c_ptr), value :: ctx
    end subroutine
    type(c_ptr) function lbf_last_error(ctx) bind(C, name="lbf_last_error")
      import
      type(c_ptr), value :: ctx
    end function
    type(c_ptr) function lbf_global_last_error() bind(C, name="lbf_global_last_error")
      import
    end function
    integer(c_int) function lbf_upload_field(ctx, id, host, dtype) bind(C, name="lbf_upload_field")
      import
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: id, dtype
    end function
    integer(c_int) function lbf_download_field(ctx, id, host, dtype) bind(C, name="lbf_download_field")
      import
      type(c_ptr), value :: ctx, host
      integer(c_int), value :: id, dtype
    end function
    integer(c_int) function lbf_upload_isfluid(ctx, isf) bind(C, name="lbf_upload_isfluid")
      import
      type(c_ptr), value :: ctx
      integer(c_int8_t), intent(in) :: isf(*)
    end function
    integer(c_int) function lbf_init_state(ctx) bind(C, name="lbf_init_state")
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lbf_step(ctx, n) bind(C, name="lbf_step")
      import
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: n
    end function
    integer(c_int) function lbf_begin_step(ctx) bind(C, name="lbf_begin_step")
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lbf_stage(ctx, id) bind(C, name="lbf_stage")
      import
      type(c_ptr), value :: ctx
      integer(c_int), value :: id
    end function
    integer(c_int) function lbf_end_step(ctx) bind(C, name="lbf_end_step")
      import
      type(c_ptr), value :: ctx
    end function
    integer(c_int) function lbf_reduce_fp64(ctx, out) bind(C, name="lbf_reduce_fp64")
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(8)
    end function
    integer(c_int) function lbf_reduce_fp64_noalloc(ctx, out) bind(C, name="lbf_reduce_fp64_noalloc")
      import
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(8)
    end function
    integer(c_int) function lbf_synchronize(ctx) bind(C, name="lbf_synchronize")
      import
      type(c_ptr), value :: ctx
    end function
  end interface

contains

  ! the message of the last failure on ctx (or of lbf_create when ctx is null)
  function lbf_error_message(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=512) :: msg
    type(c_ptr) :: cp
    character(kind=c_char), pointer :: chars(:)
    integer :: i
    if (c_associated(ctx)) then
      cp = lbf_last_error(ctx)
    else
      cp = lbf_global_last_error()
    end if
    msg = ' '
    if (.not. c_associated(cp)) return
    call c_f_pointer(cp, chars, [512])
    do i = 1, 512
      if (chars(i) == c_null_char) exit
      msg(i:i) = chars(i)
    end do
  end function

end module lbfast_hip
