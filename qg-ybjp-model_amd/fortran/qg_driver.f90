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
! Input and output follow the reference's netCDF schema (IO_ncf.f90, through
! qg_ncio): initial conditions psi000.in.nc / la000.in.nc when init_ncf_psi /
! init_ncf_la = 1, dumps N2.out.nc, psiNNN.out.nc and laNNN.out.nc.  The files
! are classic netCDF (CDF-2) rather than the reference's netCDF-4, since no
! netCDF library is available; dimensions, variables and values are the same.
! Differences by design: parameters come from the namelist /qg/ instead of a
! compile-time module; without netCDF inputs the initial conditions are raw
! float64 files (ic_prefix) or analytic fields.
program qg_driver
  use, intrinsic :: iso_c_binding
  use qg_run_params
  use qgybj_c
  use qg_ncio
  implicit none

  integer :: iktx, ikty, ix, iy, iz, it, nxt, u, ios, dump
  integer :: hunit(10), cunit, gunit, nvave
  logical :: conv
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
  write(*, '(a,4es25.17)') '# U_scale Uw_scale L_scale N0 ', U_scale, Uw_scale, L_scale, N0
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
  ! the netCDF initial conditions replace these fields (main_waqg.f90:137-139)
  if (init_ncf_psi == 1) call read_ncf_psi()
  if (init_ncf_la == 1) call read_ncf_la()
  if (grad_steps > 0) then
    call gradient_mode()
    rc = qg_destroy(ctx)
    stop
  end if
  call check(qg_set_state_physical(ctx, c_loc(psir), c_loc(BRr), c_loc(BIr), int(n1, c_int32_t)), 'set_state')
  deallocate(psir, BRr, BIr)

  dump = 0
  if (hspec_every > 0) call open_spectra()
  if (diag_every > 0) call report(0)
  if (hspec_every > 0) call spectra()
  if (conv_every > 0) then
    open(newunit=cunit, file=trim(out_prefix)//'conv.dat', action='write', status='replace')
    ! before any step the reference's tendency arrays are zero (main_waqg.f90:190)
    write(cunit, fmt=*) 0.0, 0.0, 0.0
  end if
  if (gamma_every > 0) open(newunit=gunit, file=trim(out_prefix)//'gamma.dat', action='write', status='replace')
  nvave = 0
  if (vave_every > 0) call vave()
  if (out_every > 0) call dump_n2()
  if (out_every > 0) call dump_fields(0)

  ! Euler first step, then leapfrog (main_waqg.f90:197-467): advance in
  ! chunks that end at the next diagnostic or output step
  it = 0
  do while (it < itermax)
    nxt = itermax
    if (diag_every > 0) nxt = min(nxt, (it/diag_every + 1)*diag_every)
    if (hspec_every > 0) nxt = min(nxt, (it/hspec_every + 1)*hspec_every)
    if (out_every > 0) nxt = min(nxt, (it/out_every + 1)*out_every)
    if (conv_every > 0) nxt = min(nxt, (it/conv_every + 1)*conv_every)
    if (gamma_every > 0) nxt = min(nxt, (it/gamma_every + 1)*gamma_every)
    if (vave_every > 0) nxt = min(nxt, (it/vave_every + 1)*vave_every)
    conv = .false.
    if (conv_every > 0) conv = mod(nxt, conv_every) == 0
    if (gamma_every > 0) conv = conv .or. mod(nxt, gamma_every) == 0
    if (conv) then
      ! the conversion terms read the last step's tendencies: keep them for that step only
      if (nxt - it > 1) call check(qg_step(ctx, int(nxt - it - 1, c_int32_t)), 'step')
      call check(qg_keep_tendencies(ctx, 1_c_int32_t), 'keep_tendencies')
      call check(qg_step(ctx, 1_c_int32_t), 'step')
      call conversion(nxt)
      call check(qg_keep_tendencies(ctx, 0_c_int32_t), 'keep_tendencies')
    else
      call check(qg_step(ctx, int(nxt - it, c_int32_t)), 'step')
    end if
    it = nxt
    if (diag_every > 0) then
      if (mod(it, diag_every) == 0) call report(it)
    end if
    if (out_every > 0) then
      if (mod(it, out_every) == 0) call dump_fields(it)
    end if
    if (hspec_every > 0) then
      if (mod(it, hspec_every) == 0) call spectra()
    end if
    if (vave_every > 0) then
      if (mod(it, vave_every) == 0) call vave()
    end if
  end do
  if (hspec_every > 0) then
    do iz = 1, 10
      close(hunit(iz))
    end do
  end if
  if (conv_every > 0) close(cunit)
  if (gamma_every > 0) close(gunit)

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

  ! psi (ncread_psi, IO_ncf.f90:223-293): dimensional FLOAT psi(x, y, z)
  ! over the grid, divided by U_scale L_scale
  subroutine read_ncf_psi()
    real(c_float), allocatable, target :: c(:, :, :)
    real(c_float) :: t
    allocate(c(n1, n1, n3))
    call nc_check(nc_read_field(init_ncf_psi_filename, 'psi', c), init_ncf_psi_filename)
    call nc_check(nc_read_time(init_ncf_psi_filename, t), init_ncf_psi_filename)
    write(*, '(a,es16.8,a)') '# psi read, file time ', t, ' s'
    psir = c/(U_scale*L_scale)
    deallocate(c)
  end subroutine read_ncf_psi

  ! L+A (ncread_la, IO_ncf.f90:298-371): LAr, LAi divided by Uw_scale
  subroutine read_ncf_la()
    real(c_float), allocatable, target :: c(:, :, :)
    real(c_float) :: t
    allocate(c(n1, n1, n3))
    call nc_check(nc_read_field(init_ncf_la_filename, 'LAr', c), init_ncf_la_filename)
    BRr = c/Uw_scale
    call nc_check(nc_read_field(init_ncf_la_filename, 'LAi', c), init_ncf_la_filename)
    BIr = c/Uw_scale
    call nc_check(nc_read_time(init_ncf_la_filename, t), init_ncf_la_filename)
    write(*, '(a,es16.8,a)') '# L+A read, file time ', t, ' s'
    deallocate(c)
  end subroutine read_ncf_la

  subroutine nc_check(code, fname)
    integer, intent(in) :: code
    character(len=*), intent(in) :: fname
    if (code /= 0) then
      write(*, '(a,a,a,a)') 'qg_driver: netCDF ', trim(fname), ': ', nc_error()
      stop 1
    end if
  end subroutine nc_check

  ! conv.dat: time, ca, cr of we_conversion times Uw_scale^2 (diagnostics.f90:1071-1074);
  ! gamma.dat: time in days, ca, cr, cd of gamma_conversion times Uw_scale^2 U_scale / L_scale
  ! (:1249-1253); REAL values as the reference writes them
  subroutine conversion(i)
    integer, intent(in) :: i
    real(c_double) :: cv(5)
    real :: t
    call check(qg_wave_conversion(ctx, cv), 'wave_conversion')
    t = real(i*delt)
    if (conv_every > 0) then
      if (mod(i, conv_every) == 0) then
        write(cunit, fmt=*) t, real(cv(1)*Uw_scale*Uw_scale), real(cv(2)*Uw_scale*Uw_scale)
        flush(cunit)
      end if
    end if
    if (gamma_every > 0) then
      if (mod(i, gamma_every) == 0) then
        write(gunit, fmt=*) t*(L_scale/U_scale)/(3600*24), &
          real(cv(3:5)*Uw_scale*Uw_scale*U_scale/L_scale)
        flush(gunit)
      end if
    end if
  end subroutine conversion

  ! WE_vave<count>.dat (diagnostics.f90:780-798): the map times Uw_scale^2, one value per line
  ! in E12.5 for x = 1 .. n1, then a blank line, for each y; the name as the reference's
  ! (A7,I3,A4) forms it
  subroutine vave()
    real(c_double), allocatable :: we(:, :)
    character(len=14) :: fname
    integer :: v, jx, jy
    allocate(we(n1, n1))
    call check(qg_wave_vave(ctx, we), 'wave_vave')
    write(fname, '(A7,I3,A4)') 'WE_vave', nvave, '.dat'
    open(newunit=v, file=trim(out_prefix)//fname, action='write', status='replace')
    do jy = 1, n1
      write(v, '(1x,E12.5,1x)') (real(we(jx, jy)*Uw_scale*Uw_scale), jx = 1, n1)
      write(v, fmt=*) '           '
    end do
    close(v)
    nvave = nvave + 1
    deallocate(we)
  end subroutine vave

  ! h0w.dat .. h9w.dat (files.f90:182-191)
  subroutine open_spectra()
    character(len=8) :: nm
    integer :: k
    do k = 1, 10
      write(nm, '(a,i0,a)') 'h', k - 1, 'w.dat'
      open(newunit=hunit(k), file=trim(out_prefix)//trim(nm), action='write', status='replace')
    end do
  end subroutine open_spectra

  ! hspec_waves at the heights of parameters_test1.f90:188 (diagnostics.f90:
  ! 625-634): per level one row per populated shell (mode, KE, PE, number of
  ! modes, as default REAL) and a blank line
  subroutine spectra()
    integer(c_int32_t) :: hs(10)
    real(c_double), allocatable :: sp(:, :, :)
    integer :: k, mode
    hs = int([1, n3/8, n3/4, 3*n3/8, n3/2, 5*n3/8, 3*n3/4, 7*n3/8, n3 - 2, n3], c_int32_t)
    hs = max(hs, 1_c_int32_t)
    allocate(sp(3, n1/2, 10))
    call check(qg_hspec_waves(ctx, hs, 10_c_int32_t, sp), 'hspec_waves')
    do k = 1, 10
      do mode = 0, n1/2 - 1
        if (sp(3, mode + 1, k) /= 0.d0) write(hunit(k), fmt=*) real(mode), real(sp(1, mode + 1, k)), &
                                                                real(sp(2, mode + 1, k)), real(sp(3, mode + 1, k))
      end do
      write(hunit(k), *) '           '
      flush(hunit(k))
    end do
    deallocate(sp)
  end subroutine spectra

  ! N^2 at the staggered levels in s^-2 (ncdump_N2, IO_ncf.f90:376-432)
  subroutine dump_n2()
    real(c_float), allocatable, target :: w(:)
    allocate(w(n3))
    w = real(N2s*N0*N0, c_float)
    call nc_check(nc_write_profile(trim(out_prefix)//'N2.out.nc', 'N2', w), trim(out_prefix)//'N2.out.nc')
    deallocate(w)
  end subroutine dump_n2

  ! psi and L+A in r-space after step i, as ncdump_psi / ncdump_la
  ! (IO_ncf.f90:15-110, 116-217): FLOAT values psi U_scale L_scale and
  ! BR, BI times Uw_scale, time i dt (default REAL, main_waqg.f90:197, 328)
  ! times L_scale/U_scale in seconds; files numbered from 000
  subroutine dump_fields(i)
    integer, intent(in) :: i
    complex(c_double_complex), allocatable, target :: pk(:, :, :), brk(:, :, :), bik(:, :, :)
    real(c_double), allocatable, target :: r(:, :, :)
    real(c_float), allocatable, target :: w(:, :, :, :)
    real :: time
    real(c_float) :: ts
    character(len=16) :: tag
    character(len=512) :: fname
    allocate(pk(iktx, ikty, n3), brk(iktx, ikty, n3), bik(iktx, ikty, n3), r(n1, n1, n3))
    call check(qg_get_state(ctx, c_loc(pk), c_null_ptr, c_null_ptr, c_loc(brk), c_loc(bik), c_null_ptr, &
                            c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr, c_null_ptr), 'get_state')
    time = i*delt
    ts = real(time*L_scale/U_scale, c_float)
    write(tag, '(i0.3)') dump
    allocate(w(n1, n1, n3, 2))
    call check(qg_fft_c2r(ctx, c_loc(pk), c_loc(r), int(n1, c_int32_t), int(n3, c_int32_t)), 'c2r')
    w(:, :, :, 1) = real(r*U_scale*L_scale, c_float)
    fname = trim(out_prefix)//'psi'//trim(tag)//'.out.nc'
    call nc_check(nc_write_fields(fname, 'psi', w(:, :, :, 1:1), ts), fname)
    call check(qg_fft_c2r(ctx, c_loc(brk), c_loc(r), int(n1, c_int32_t), int(n3, c_int32_t)), 'c2r')
    w(:, :, :, 1) = real(r*Uw_scale, c_float)
    call check(qg_fft_c2r(ctx, c_loc(bik), c_loc(r), int(n1, c_int32_t), int(n3, c_int32_t)), 'c2r')
    w(:, :, :, 2) = real(r*Uw_scale, c_float)
    fname = trim(out_prefix)//'la'//trim(tag)//'.out.nc'
    call nc_check(nc_write_fields(fname, 'LAr,LAi', w, ts), fname)
    deallocate(pk, brk, bik, r, w)
    dump = dump + 1
  end subroutine dump_fields

end program qg_driver
