! Run parameters of the Fortran driver and the tables derived from them.
!
! The reference fixes its parameters at compile time (parameters_test1.f90);
! this driver reads the same free parameters from a namelist at run time and
! derives the rest with the reference's expressions, in the precision the
! reference evaluates them (default-REAL literals and the REAL(4) powers of
! the integrating factor, main_waqg.f90:371-372).
module qg_run_params
  use, intrinsic :: iso_c_binding, only: c_double, c_int32_t
  implicit none

  double precision, parameter :: twopi = 4.d0 * asin(1.d0)
  double precision, parameter :: L1 = twopi, L3 = twopi

  integer, parameter :: constant_N = 1, skewed_gaussian = 2, exponential_N = 3

  ! free parameters (namelist /qg/)
  integer :: n1 = 64, n3 = 32
  double precision :: delt = 1.d-3
  real :: gamma = 1e-3                 ! REAL(4) literal, as in the reference
  double precision :: Ro = 0.1d0, Bu = 1.d0, W2F = 1.d-2
  double precision :: coeff1 = 0.d0, coeff2 = 10.d0, coeff1w = 0.d0, coeff2w = 10.d0, coeffz = 0.d0
  integer :: ilap1 = 4, ilap2 = 4, ilap1w = 4, ilap2w = 4
  integer :: fixed_flow = 0, no_feedback = 0, passive_scalar = 0, linear = 0, no_dispersion = 0, inviscid = 0
  integer :: ybj_plus = 1              ! 0: the regular YBJ operator (sumB, compute_sigma, compute_A)
  integer :: stratification = constant_N
  double precision :: n2_ratio = 100.d0
  integer :: itermax = 10              ! steps to run (the first one is the Euler step)
  integer :: diag_every = 1            ! print the energy diagnostics every diag_every steps (0: never)
  integer :: out_every = 0             ! netCDF dumps of psi and L+A in r-space every out_every steps (0: never)
  character(len=256) :: ic_prefix = ' '   ! read <prefix>psi.in, BR.in, BI.in (float64 n1*n1*n3); blank: analytic
  character(len=256) :: out_prefix = ' '  ! prefix of the dumps; final spectral state to <prefix>state.bin unless blank
  ! gradient mode (no counterpart in the reference): grad_steps > 0 computes, instead of the run,
  ! the deep wave energy after grad_steps steps and its gradient with respect to the initial psi
  ! (checkpoints every grad_every steps, levels with z_s < grad_zdeep); the gradient goes to
  ! <out_prefix>grad.bin (float64 n1*n1*n3)
  integer :: grad_steps = 0, grad_every = 20
  double precision :: grad_zdeep = 0.8d0*6.28318530717958647692d0
  ! netCDF input and output (IO_ncf.f90; main_waqg.f90:137-139, 153-155, 470-471): with
  ! init_ncf_psi = 1 (init_ncf_la = 1) psi (L+A) is read from init_ncf_psi_filename
  ! (init_ncf_la_filename) in place of the fields above; with out_every > 0, N2.out.nc once and
  ! psiNNN.out.nc, laNNN.out.nc at step 0 and every out_every steps (names prefixed with
  ! out_prefix).  Values are dimensional with the reference's scales (parameters_test1.f90:8,
  ! 45-50: default-REAL literals as there); they scale the files only, the step takes Ro, Bu and
  ! W2F as given.
  ! wave spectra (hspec_waves, diagnostics.f90:547-638) at the reference's 10 heights
  ! (parameters_test1.f90:188) to <out_prefix>h0w.dat .. h9w.dat at step 0 and every hspec_every
  ! steps (0: never)
  integer :: hspec_every = 0
  ! wave energy conversion terms to <out_prefix>conv.dat (we_conversion, diagnostics.f90:967-1076;
  ! at step 0 and every conv_every steps) and <out_prefix>gamma.dat (gamma_conversion,
  ! :1080-1256; every gamma_every steps), as the reference writes them (files.f90:140-141)
  integer :: conv_every = 0, gamma_every = 0
  ! vertically averaged wave energy density maps (we_vave, diagnostics.f90:750-813) to
  ! <out_prefix>WE_vave<n>.dat at step 0 and every vave_every steps (0: never)
  integer :: vave_every = 0
  integer :: init_ncf_psi = 0, init_ncf_la = 0
  character(len=256) :: init_ncf_psi_filename = 'psi000.in.nc', init_ncf_la_filename = 'la000.in.nc'
  double precision :: U_scale = 0.25, Uw_scale = 2.5e-5
  double precision :: L_scale = 6.28318530718*5e4/L1     ! dom_x/L1
  double precision :: N0 = (25./8.)*twopi*1e-4            ! (25/8) 2 pi cor

  namelist /qg/ n1, n3, delt, gamma, Ro, Bu, W2F, coeff1, coeff2, coeff1w, coeff2w, coeffz, &
                ilap1, ilap2, ilap1w, ilap2w, fixed_flow, no_feedback, passive_scalar, linear, &
                no_dispersion, inviscid, ybj_plus, stratification, n2_ratio, itermax, diag_every, out_every, &
                ic_prefix, out_prefix, grad_steps, grad_every, grad_zdeep, init_ncf_psi, init_ncf_la, &
                init_ncf_psi_filename, init_ncf_la_filename, U_scale, Uw_scale, L_scale, N0, hspec_every, &
                conv_every, gamma_every, vave_every

contains

  ! coeff*(64./(1.*n1))**(4./3.)*(3./n1)**(2*(ilap-1))  (parameters_test1.f90:106-109)
  double precision function nuh(coeff, ilap)
    double precision, intent(in) :: coeff
    integer, intent(in) :: ilap
    nuh = coeff * (64./(1.*n1))**(4./3.) * (3./n1)**(2*(ilap-1))
  end function nuh

  ! coeffz*(64./(1.*n1))**(4./3.)  (parameters_test1.f90:113)
  double precision function nuz()
    nuz = coeffz * (64./(1.*n1))**(4./3.)
  end function nuz

  ! Bit set of include/qgybj.h's QG_* flags
  integer(c_int32_t) function flag_bits()
    flag_bits = 0
    if (fixed_flow == 1) flag_bits = ior(flag_bits, 1)
    if (no_feedback == 1) flag_bits = ior(flag_bits, 2)
    if (passive_scalar == 1) flag_bits = ior(flag_bits, 4)
    if (linear == 1) flag_bits = ior(flag_bits, 8)
    if (no_dispersion == 1) flag_bits = ior(flag_bits, 16)
    if (inviscid == 1) flag_bits = ior(flag_bits, 32)
    if (ybj_plus == 0) flag_bits = ior(flag_bits, 64)
  end function flag_bits

  ! kxa, kya of init.f90:32-41 (the Nyquist row is relabelled 0)
  integer function kya_of(iky)
    integer, intent(in) :: iky
    integer :: j
    j = iky - 1
    if (j < n1/2) then
      kya_of = j
    else if (j == n1/2) then
      kya_of = 0
    else
      kya_of = j - n1
    end if
  end function kya_of

  ! int_factor(kx,ky) of main_waqg.f90:371 for one pair of operators:
  ! delt*(nu1*((1.*kx)**(2.*ilapa)+(1.*ky)**(2.*ilapa)) + nu2*(...))
  subroutine int_factor_table(nu1, ilapa, nu2, ilapb, tab)
    double precision, intent(in) :: nu1, nu2
    integer, intent(in) :: ilapa, ilapb
    real(c_double), intent(out) :: tab(:, :)          ! (iktx, ikty)
    integer :: ikx, iky, kx, ky
    do iky = 1, n1
      ky = kya_of(iky)
      do ikx = 1, n1/2 + 1
        kx = ikx - 1
        tab(ikx, iky) = delt * ( nu1*((1.*kx)**(2.*ilapa) + (1.*ky)**(2.*ilapa)) &
                               + nu2*((1.*kx)**(2.*ilapb) + (1.*ky)**(2.*ilapb)) )
      end do
    end do
  end subroutine int_factor_table

  ! N^2(z): constant and skewed Gaussian as init.f90:186-195 (with the
  ! default-REAL constants of parameters_test1.f90:58-63); exponential with
  ! N^2 = 1 at the surface z = L3 and 1/n2_ratio at the bottom z = 0
  double precision function n2_of(z)
    double precision, intent(in) :: z
    double precision, parameter :: N02_sg = 0.537713935783168
    double precision, parameter :: N12_sg = 2.684198470106461
    double precision, parameter :: sigma_sg = 0.648457170048730
    double precision, parameter :: z0_sg = 6.121537923499139
    double precision, parameter :: alpha_sg = -5.338431587899242
    select case (stratification)
    case (constant_N)
      n2_of = 1.d0
    case (skewed_gaussian)
      n2_of = N12_sg*exp(-((z-z0_sg)**2)/(sigma_sg**2))*(1.+erf(alpha_sg*(z-z0_sg)/(sigma_sg*sqrt(2.)))) + N02_sg
    case (exponential_N)
      n2_of = exp(log(n2_ratio) * (z - L3) / L3)
    case default
      stop "undefined stratification profile"
    end select
  end function n2_of

end module qg_run_params
