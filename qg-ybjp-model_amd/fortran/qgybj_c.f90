! ISO_C_BINDING interface to the MI355X QG-YBJ+ step (include/qgybj.h).
!
! The reference (oliasselin/QG-YBJp-model) runs its loop body as Fortran
! calls over static arrays (main_waqg.f90:326-467); a Fortran driver calls
! these C entry points instead.  Array arguments are passed as c_ptr
! (c_loc of a TARGET array, or c_null_ptr where the C side accepts NULL):
! complex(c_double_complex) X(iktx, ikty, n3) spectral fields in the
! reference's layout and real(c_double) X(ldx, n2, nz) physical fields.
module qgybj_c
  use, intrinsic :: iso_c_binding
  implicit none

  integer(c_int), parameter :: QG_OK = 0, QG_ERR_ARG = 1, QG_ERR_HIP = 2, QG_ERR_NOMEM = 3, &
                               QG_ERR_NONFINITE = 4, QG_ERR_STATE = 5
  integer(c_int32_t), parameter :: QG_FIXED_FLOW = 1, QG_NO_FEEDBACK = 2, QG_PASSIVE_SCALAR = 4, &
                                   QG_LINEAR = 8, QG_NO_DISPERSION = 16, QG_INVISCID = 32

  ! struct qg_params
  type, bind(C) :: qg_params_t
    integer(c_int32_t) :: n1, n2, n3, flags
    real(c_double) :: dt, gamma, Ro, Bu, W2F, nuz
    type(c_ptr) :: ifac_q, ifac_w, N2_u, N2_s
  end type qg_params_t

  interface
    integer(c_int) function qg_create(p, ctx) bind(C, name="qg_create")
      import :: c_int, c_ptr, qg_params_t
      type(qg_params_t), intent(in) :: p
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function qg_destroy(ctx) bind(C, name="qg_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    type(c_ptr) function qg_last_error(ctx) bind(C, name="qg_last_error")
      import :: c_ptr
      type(c_ptr), value :: ctx
    end function

    integer(c_int) function qg_set_state(ctx, psik, qk, qok, BRk, BIk, BRok, BIok, ARk, AIk, iter) &
        bind(C, name="qg_set_state")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, psik, qk, qok, BRk, BIk, BRok, BIok, ARk, AIk
      integer(c_int32_t), value :: iter
    end function

    integer(c_int) function qg_set_state_physical(ctx, psir, BRr, BIr, ldx) bind(C, name="qg_set_state_physical")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, psir, BRr, BIr
      integer(c_int32_t), value :: ldx
    end function

    integer(c_int) function qg_get_state(ctx, psik, qk, qok, BRk, BIk, BRok, BIok, ARk, AIk, CRk, CIk) &
        bind(C, name="qg_get_state")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, psik, qk, qok, BRk, BIk, BRok, BIok, ARk, AIk, CRk, CIk
    end function

    integer(c_int) function qg_step(ctx, nsteps) bind(C, name="qg_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: nsteps
    end function

    integer(c_int) function qg_iter(ctx, iter) bind(C, name="qg_iter")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: iter
    end function

    integer(c_int) function qg_diag(ctx, out) bind(C, name="qg_diag")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(6)
    end function

    integer(c_int) function qg_fft_c2r(ctx, k, r, ldx, nz) bind(C, name="qg_fft_c2r")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, k, r
      integer(c_int32_t), value :: ldx, nz
    end function

    integer(c_int) function qg_fft_r2c(ctx, r, ldx, k, nz) bind(C, name="qg_fft_r2c")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, r, k
      integer(c_int32_t), value :: ldx, nz
    end function
  end interface

contains

  ! message of the last failure on a context, as a Fortran string
  function qg_error_string(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    character(kind=c_char), pointer :: cs(:)
    type(c_ptr) :: p
    integer :: i
    msg = ' '
    p = qg_last_error(ctx)
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [256])
    do i = 1, 256
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    end do
  end function qg_error_string

end module qgybj_c
