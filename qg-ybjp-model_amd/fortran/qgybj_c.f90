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
                                   QG_LINEAR = 8, QG_NO_DISPERSION = 16, QG_INVISCID = 32, &
                                   QG_YBJ_REGULAR = 64

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

    ! wave spectra at nlev 1-based levels: out(3, n1/2, nlev) = KE, PE, modes per shell
    integer(c_int) function qg_hspec_waves(ctx, levels, nlev, out) bind(C, name="qg_hspec_waves")
      import :: c_int, c_int32_t, c_ptr, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: levels(*)
      integer(c_int32_t), value :: nlev
      real(c_double), intent(out) :: out(*)
    end function

    ! vertically averaged wave energy density 0.5 (BR^2 + BI^2) / n3, out(n1, n2) per member
    integer(c_int) function qg_wave_vave(ctx, out) bind(C, name="qg_wave_vave")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(*)
    end function

    ! keep each step's wave tendencies nB, rB (on /= 0) for the conversion terms
    integer(c_int) function qg_keep_tendencies(ctx, on) bind(C, name="qg_keep_tendencies")
      import :: c_int, c_int32_t, c_ptr
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: on
    end function

    ! we_conversion ca, cr and gamma_conversion ca, cr, cd per member, nondimensional
    integer(c_int) function qg_wave_conversion(ctx, out) bind(C, name="qg_wave_conversion")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: ctx
      real(c_double), intent(out) :: out(*)
    end function

    ! QG vertical velocity w of the current psi (omega equation), shaped like a state field
    integer(c_int) function qg_omega(ctx, wk) bind(C, name="qg_omega")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx, wk
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

    integer(c_int) function qg_sync(ctx) bind(C, name="qg_sync")
      import :: c_int, c_ptr
      type(c_ptr), value :: ctx
    end function

    ! batches: `members` members with the same grid and parameters in one
    ! context; every host array gains a trailing member dimension
    integer(c_int) function qg_create_batch(p, members, ctx) bind(C, name="qg_create_batch")
      import :: c_int, c_int32_t, c_ptr, qg_params_t
      type(qg_params_t), intent(in) :: p
      integer(c_int32_t), value :: members
      type(c_ptr), intent(out) :: ctx
    end function

    integer(c_int) function qg_members(ctx, members) bind(C, name="qg_members")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: members
    end function

    ! kernel variants (QG_NVARIANTS values) and the tuner
    integer(c_int) function qg_set_variants(ctx, v, n) bind(C, name="qg_set_variants")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(in) :: v(*)
      integer(c_int32_t), value :: n
    end function

    integer(c_int) function qg_get_variants(ctx, v, n) bind(C, name="qg_get_variants")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx
      integer(c_int32_t), intent(out) :: v(*)
      integer(c_int32_t), value :: n
    end function

    integer(c_int) function qg_tune(ctx, steps, best, best_ms) bind(C, name="qg_tune")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ctx, best, best_ms
      integer(c_int32_t), value :: steps
    end function

    ! ensembles of contexts advanced together (graph replays)
    integer(c_int) function qg_ensemble_create(members, n, ens) bind(C, name="qg_ensemble_create")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), intent(in) :: members(*)
      integer(c_int32_t), value :: n
      type(c_ptr), intent(out) :: ens
    end function

    integer(c_int) function qg_ensemble_destroy(ens) bind(C, name="qg_ensemble_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: ens
    end function

    type(c_ptr) function qg_ensemble_error(ens) bind(C, name="qg_ensemble_error")
      import :: c_ptr
      type(c_ptr), value :: ens
    end function

    integer(c_int) function qg_ensemble_set_graph_steps(ens, k) bind(C, name="qg_ensemble_set_graph_steps")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ens
      integer(c_int32_t), value :: k
    end function

    integer(c_int) function qg_ensemble_step(ens, nsteps) bind(C, name="qg_ensemble_step")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: ens
      integer(c_int32_t), value :: nsteps
    end function

    ! tangent-linear and adjoint (gradients with respect to the initial psi);
    ! controls and gradients are physical real(c_double) X(ldx, n2, n3)
    integer(c_int) function qg_adj_create(ctx, checkpoint_every, z_deep, adj) bind(C, name="qg_adj_create")
      import :: c_int, c_ptr, c_int32_t, c_double
      type(c_ptr), value :: ctx
      integer(c_int32_t), value :: checkpoint_every
      real(c_double), value :: z_deep
      type(c_ptr), intent(out) :: adj
    end function

    integer(c_int) function qg_adj_destroy(adj) bind(C, name="qg_adj_destroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: adj
    end function

    type(c_ptr) function qg_adj_error(adj) bind(C, name="qg_adj_error")
      import :: c_ptr
      type(c_ptr), value :: adj
    end function

    integer(c_int) function qg_adj_set_fused(adj, fused) bind(C, name="qg_adj_set_fused")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: adj
      integer(c_int32_t), value :: fused
    end function

    integer(c_int) function qg_adj_objective(adj, psir, BRr, BIr, ldx, nsteps, J) bind(C, name="qg_adj_objective")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: adj, psir, BRr, BIr, J
      integer(c_int32_t), value :: ldx, nsteps
    end function

    integer(c_int) function qg_adj_tlm(adj, psir, BRr, BIr, ldx, nsteps, dpsir, dJ, dstate) bind(C, name="qg_adj_tlm")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: adj, psir, BRr, BIr, dpsir, dJ, dstate
      integer(c_int32_t), value :: ldx, nsteps
    end function

    integer(c_int) function qg_adj_apply(adj, psir, BRr, BIr, ldx, nsteps, state_bar, psi_bar) &
        bind(C, name="qg_adj_apply")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: adj, psir, BRr, BIr, state_bar, psi_bar
      integer(c_int32_t), value :: ldx, nsteps
    end function

    integer(c_int) function qg_adj_gradient(adj, psir, BRr, BIr, ldx, nsteps, J, grad) bind(C, name="qg_adj_gradient")
      import :: c_int, c_ptr, c_int32_t
      type(c_ptr), value :: adj, psir, BRr, BIr, J, grad
      integer(c_int32_t), value :: ldx, nsteps
    end function
  end interface

contains

  ! a NUL-terminated C string (qg_last_error, qg_ensemble_error,
  ! qg_adj_error) as a Fortran string
  function c_string(p) result(msg)
    type(c_ptr), intent(in) :: p
    character(len=256) :: msg
    character(kind=c_char), pointer :: cs(:)
    integer :: i
    msg = ' '
    if (.not. c_associated(p)) return
    call c_f_pointer(p, cs, [256])
    do i = 1, 256
      if (cs(i) == c_null_char) exit
      msg(i:i) = cs(i)
    end do
  end function c_string

  ! message of the last failure on a context, as a Fortran string
  function qg_error_string(ctx) result(msg)
    type(c_ptr), intent(in) :: ctx
    character(len=256) :: msg
    msg = c_string(qg_last_error(ctx))
  end function qg_error_string

end module qgybj_c
