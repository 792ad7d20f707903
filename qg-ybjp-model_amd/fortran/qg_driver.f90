! Fortran integration driver over the C ABI (include/qgybj.h).
!
!   qg_driver [namelist-file]
!
! Same loop structure as the reference's main program (main_waqg.f90):
! set up the grid and base state, load psi and L+A in r-space, run the Euler
! first step and the leapfrog steps, print energy diagnostics and dump r-space
! fields at fixed intervals.  The loop body runs on the GPU through the C ABI;
! this program only forms the parameters (qg_run_params), moves the initial
! and output fields and reports.
!
! Differences from the reference, by design: parameters come from the
! namelist /qg/ instead of a compile-time module; initial conditions are raw
! float64 files or the analytic fields of generate_fields_stag
! (init.f90:795-882) instead of netCDF; r-space dumps are raw float32 files
! (<out_prefix>psiNNNN.bin, laNNNN.bin with BR then BI) in the layout of the
! reference's netCDF variables (IO_ncf.f90:24-120), since no netCDF library is
! available.
program qg_driver
  use, intrinsic :: iso_c_binding
  use qg_run_params
  use qgybj_c
  implicit none

  integer :: iktx, ikty, ix, iy, iz, it, nxt, u, ios, dump
  integer(c_int) :: rc
  integer(c_int32_t) :: iter
  character(len=512) :: arg
  type(qg_params_t) :: p
  type(c_ptr) :: ctx
  real(c_double), allocatable, target :: ifq(:, :), ifw(:, :), N2u(:), N2s(:)
  real(c_double), allocatable, target :: psir(:, :, :), BRr(:, :, :), BIr(:, :, :)
  complex(c_double_complex), allocatable, target :: fk(:, :, :, :)
  real(c_double) :: d(6)
  double precision :: x, y, zs, dz, nuh1, nuh2, nuh1w, nuh2w

  if (command_argument_count() >= 1) then
    call get_command_argument(1, arg)
    open(newunit=u, file=trim(arg), status='old', action='read', iostat=ios)
    if (ios /= 0) stop "qg_driver: cannot open the namelist file"
    read(u, nml=qg, iostat=ios)
    if (ios /= 0) stop "qg_driver: cannot read namelist /qg/"
    close(u)
  end if

  iktx = n1/2 + 1
  ikty = n1
  dz = L3/n3

  ! base state and integrating-factor tables (init.f90:177-264, main_waqg.f90:371-372)
  nuh1 = nuh(coeff1, ilap1)
  nuh2 = nuh(coeff2, ilap2)
  nuh1w = nuh(coeff1w, ilap1w)
  nuh2w = nuh(coeff2w, ilap2w)
  allocate(ifq(iktx, ikty), ifw(iktx, ikty), N2u(n3), N2s(n3))
  call int_factor_table(nuh1, ilap1, nuh2, ilap2, ifq)
  call int_factor_table(nuh1w, ilap1w, nuh2w, ilap2w, ifw)
  do iz = 1, n3
    N2u(iz) = n2_of(L3*iz/n3)
    N2s(iz) = n2_of(L3*iz/n3 - dz/2)
  end do

  p%n1 = n1
  p%n2 = n1
  p%n3 = n3
  p%flags = flag_bits()
  p%dt = delt
  p%gamma = gamma
  p%Ro = Ro
  p%Bu = Bu
  p%W2F = W2F
  p%nuz = nuz()
  p%ifac_q = c_loc(ifq)
  p%ifac_w = c_loc(ifw)
  p%N2_u = c_loc(N2u)
  p%N2_s = c_loc(N2s)
  write(*, '(a,5es25.17)') '# nuh1 nuh2 nuh1w nuh2w nuz ', nuh1, nuh2, nuh1w, nuh2w, p%nuz
  flush(6)
  rc = qg_create(p, ctx)
  if (rc /= QG_OK) then
    write(*, '(a,i0)') 'qg_driver: qg_create failed, code ', rc
    stop 1
  end if

  ! initial conditions in r-space
  allocate(psir(n1, n1, n3), BRr(n1, n1, n3), BIr(n1, n1, n3))
  if (len_trim(ic_prefix) == 0) then
    do iz = 1, n3
      zs = L3*iz/n3 - dz/2
      do iy = 1, n1
        y = L1*(iy - 1)/n1
        do ix = 1, n1
          x = L1*(ix - 1)/n1
          psir(ix, iy, iz) = sin(x)*sin(y)
          BRr(ix, iy, iz) = cos(8*zs/2.)
          BIr(ix, iy, iz) = 0.d0
        end do
      end do
    end do
  else
    call read_field(trim(ic_prefix)//'psi.in', psir)
    call read_field(trim(ic_prefix)//'BR.in', BRr)
    call read_field(trim(ic_prefix)//'BI.in', BIr)
  end if
  if (grad_steps > 0) then
    call gradient_mode()
    rc = qg_destroy(ctx)
    stop
  end if
  call check(qg_set_state_physical(ctx, c_loc(psir), c_loc(BRr), c_loc(BIr), int(n1, c_int32_t)), 'set_state')
  deallocate(psir, BRr, BIr)

  dump = 0
  if (diag_every > 0) call report(0)
  if (out_every > 0) call dump_fields()

  ! Euler first step, then leapfrog (main_waqg.f90:197-467): advance in
  ! chunks that end at the next diagnostic or output step
  it = 0
  do while (it < itermax)
    nxt = itermax
    if (diag_every > 0) nxt = min(nxt, (it/diag_every + 1)*diag_every)
    if (out_every > 0) nxt = min(nxt, (it/out_every + 1)*out_every)
    call check(qg_step(ctx, int(nxt - it, c_int32_t)), 'step')
    it = nxt
    if (diag_every > 0) then
      if (mod(it, diag_every) == 0) call report(it)
    end if
    if (out_every > 0) then
      if (mod(it, out_every) == 0) call dump_fields()
    end if
  end do

  if (len_trim(out_prefix) > 0) then
    allocate(fk(iktx, ikty, n3, 9))
    call check(qg_get_state(ctx, c_loc(fk(1,1,1,1)), c_loc(fk(1,1,1,2)), c_loc(fk(1,1,1,3)), &
                            c_loc(fk(1,1,1,4)), c_loc(fk(1,1,1,5)), c_loc(fk(1,1,1,6)), c_loc(fk(1,1,1,7)), &
                            c_loc(fk(1,1,1,8)), c_loc(fk(1,1,1,9)), c_null_ptr, c_null_ptr), 'get_state')
    open(newunit=u, file=trim(out_prefix)//'state.bin', access='stream', form='unformatted', &
         status='replace', action='write')
    write(u) fk
    close(u)
    deallocate(fk)
  end if

  call check(qg_iter(ctx, iter), 'iter')
  write(*, '(a,i0)') '# done iter ', iter
  rc = qg_destroy(ctx)

contains

  ! J (deep wave energy after grad_steps steps) and dJ/dpsi0 through the
  ! adjoint entry points, from the initial conditions in psir, BRr, BIr
  subroutine gradient_mode()
    type(c_ptr) :: adj
    real(c_double), target :: J
    real(c_double), allocatable, target :: g(:, :, :)
    integer :: v, e
    e = qg_adj_create(ctx, int(grad_every, c_int32_t), grad_zdeep, adj)
    if (e /= QG_OK) then
      write(*, '(a,i0,a,a)') 'qg_driver: qg_adj_create failed, code ', e, ': ', trim(qg_error_string(ctx))
      stop 1
    end if
    allocate(g(n1, n1, n3))
    e = qg_adj_gradient(adj, c_loc(psir), c_loc(BRr), c_loc(BIr), int(n1, c_int32_t), int(grad_steps, c_int32_t), &
                        c_loc(J), c_loc(g))
    if (e /= QG_OK) then
      write(*, '(a,i0,a,a)') 'qg_driver: qg_adj_gradient failed, code ', e, ': ', trim(c_string(qg_adj_error(adj)))
      stop 1
    end if
    write(*, '(a,i0,a,es25.17)') '# grad steps ', grad_steps, ' J ', J
    write(*, '(a,es25.17)') '# grad norm ', sqrt(sum(g*g))
    if (len_trim(out_prefix) > 0) then
      open(newunit=v, file=trim(out_prefix)//'grad.bin', access='stream', form='unformatted', &
           status='replace', action='write')
      write(v) g
      close(v)
    end if
    e = qg_adj_destroy(adj)
    deallocate(g)
  end subroutine gradient_mode

  subroutine check(code, what)
    integer(c_int), intent(in) :: code
    character(len=*), intent(in) :: what
    if (code /= QG_OK) then
      write(*, '(a,a,a,i0,a,a)') 'qg_driver: ', what, ' failed, code ', code, ': ', trim(qg_error_string(ctx))
      stop 1
    end if
  end subroutine check

  subroutine read_field(fname, f)
    character(len=*), intent(in) :: fname
    real(c_double), intent(out) :: f(:, :, :)
    integer :: v, e
    open(newunit=v, file=fname, access='stream', form='unformatted', status='old', action='read', iostat=e)
    if (e /= 0) then
      write(*, '(a,a)') 'qg_driver: cannot open ', fname
      stop 1
    end if
    read(v) f
    close(v)
  end subroutine read_field

  ! one line per report: iter, time, KE, PE, WKE, WPE, WCE, enstrophy
  subroutine report(i)
    integer, intent(in) :: i
    call check(qg_diag(ctx, d), 'diag')
    write(*, '(i8,es16.8,6es25.16)') i, i*delt, d
  end subroutine report

  ! psi and L+A in r-space, float32, as the reference's netCDF dumps
  subroutine dump_fields()
    complex(c_double_complex), allocatable, target :: pk(:, :, :), brk(:, :, :), bik(:, :, :)
    real(c_double), allocatable, target :: r(:, :, :)
    character(len=8) :: tag
    integer :: v
    if (len_trim(out_prefix) == 0) return
    allocate(pk(iktx, ikty, n3), brk(iktx, ikty, n3), bik(iktx, ikty, n3), r(n1, n1, n3))
    call check(qg_get_state(ctx, c_loc(pk), c_null_ptr, c_null_ptr, c_loc(brk), c_loc(bik), c_null_ptr, &
                            c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), 'get_state')
    write(tag, '(i4.4)') dump
    open(newunit=v, file=trim(out_prefix)//'psi'//trim(tag)//'.bin', access='stream', form='unformatted', &
         status='replace', action='write')
    call check(qg_fft_c2r(ctx, c_loc(pk), c_loc(r), int(n1, c_int32_t), int(n3, c_int32_t)), 'c2r')
    write(v) real(r, 4)
    close(v)
    open(newunit=v, file=trim(out_prefix)//'la'//trim(tag)//'.bin', access='stream', form='unformatted', &
         status='replace', action='write')
    call check(qg_fft_c2r(ctx, c_loc(brk), c_loc(r), int(n1, c_int32_t), int(n3, c_int32_t)), 'c2r')
    write(v) real(r, 4)
    call check(qg_fft_c2r(ctx, c_loc(bik), c_loc(r), int(n1, c_int32_t), int(n3, c_int32_t)), 'c2r')
    write(v) real(r, 4)
    close(v)
    deallocate(pk, brk, bik, r)
    dump = dump + 1
  end subroutine dump_fields

end program qg_driver
