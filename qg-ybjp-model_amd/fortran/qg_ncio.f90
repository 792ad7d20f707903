! netCDF dumps and initial conditions of the driver, in the reference's
! schema (IO_ncf.f90): r-space fields over the dimensions x, y, z (Fortran
! order; z, y, x in the file), a one-element time over t, FLOAT values in
! dimensional units.  The files are classic netCDF (CDF-2), written and read
! by qg_ncio.c.
module qg_ncio
  use, intrinsic :: iso_c_binding
  implicit none
  private
  public :: nc_write_fields, nc_write_profile, nc_read_field, nc_read_time, nc_error

  interface
    integer(c_int) function qg_nc_write(path, dims, dim_len, ndims, vars, var_ndims, var_dims, var_data, nvars) &
        bind(C, name='qg_nc_write')
      import :: c_int, c_int32_t, c_char, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: path, dims, vars
      integer(c_int32_t), dimension(*), intent(in) :: dim_len, var_ndims, var_dims
      integer(c_int32_t), value :: ndims, nvars
      type(c_ptr), dimension(*), intent(in) :: var_data
    end function qg_nc_write
    integer(c_int) function qg_nc_read(path, var, dims, dim_len, ndims, out) bind(C, name='qg_nc_read')
      import :: c_int, c_int32_t, c_char, c_ptr
      character(kind=c_char), dimension(*), intent(in) :: path, var, dims
      integer(c_int32_t), dimension(*), intent(in) :: dim_len
      integer(c_int32_t), value :: ndims
      type(c_ptr), value :: out
    end function qg_nc_read
    type(c_ptr) function qg_nc_error_c() bind(C, name='qg_nc_error')
      import :: c_ptr
    end function qg_nc_error_c
    integer(c_size_t) function c_strlen(s) bind(C, name='strlen')
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
    end function c_strlen
  end interface

contains

  ! Last error of the netCDF routines as a Fortran string
  function nc_error() result(s)
    character(len=:), allocatable :: s
    type(c_ptr) :: p
    character(kind=c_char), pointer :: c(:)
    integer :: i, n
    p = qg_nc_error_c()
    n = int(c_strlen(p))
    call c_f_pointer(p, c, [n])
    allocate(character(len=n) :: s)
    do i = 1, n
      s(i:i) = c(i)
    end do
  end function nc_error

  ! Fields f(:, :, :, k) named by the comma-separated `names`, and `time`
  ! (IO_ncf.f90:66-101 for psi, :171-208 for LAr, LAi): dimensions x, y, z,
  ! t = 1, in that order, variables in the order given, then time.
  integer function nc_write_fields(fname, names, f, time) result(rc)
    character(len=*), intent(in) :: fname, names
    real(c_float), intent(in), target, contiguous :: f(:, :, :, :)
    real(c_float), intent(in) :: time
    real(c_float), target :: t(1)
    integer(c_int32_t) :: nd(size(f, 4) + 1), vd(3*size(f, 4) + 1), k
    type(c_ptr) :: ptr(size(f, 4) + 1)
    t(1) = time
    do k = 1, size(f, 4)
      nd(k) = 3
      vd(3*k - 2:3*k) = [2, 1, 0]      ! z, y, x: the file's slowest-first order of f(x, y, z)
      ptr(k) = c_loc(f(1, 1, 1, k))
    end do
    nd(size(f, 4) + 1) = 1
    vd(3*size(f, 4) + 1) = 3           ! t
    ptr(size(f, 4) + 1) = c_loc(t)
    rc = qg_nc_write(trim(fname)//c_null_char, 'x,y,z,t'//c_null_char, &
                     [int(size(f, 1), c_int32_t), int(size(f, 2), c_int32_t), int(size(f, 3), c_int32_t), 1_c_int32_t], &
                     4_c_int32_t, trim(names)//',time'//c_null_char, nd, vd, ptr, int(size(f, 4) + 1, c_int32_t))
  end function nc_write_fields

  ! One profile p(z) named `name` over the dimension z (IO_ncf.f90:394-424)
  integer function nc_write_profile(fname, name, p) result(rc)
    character(len=*), intent(in) :: fname, name
    real(c_float), intent(in), target, contiguous :: p(:)
    type(c_ptr) :: ptr(1)
    ptr(1) = c_loc(p)
    rc = qg_nc_write(trim(fname)//c_null_char, 'z'//c_null_char, [int(size(p), c_int32_t)], 1_c_int32_t, &
                     trim(name)//c_null_char, [1_c_int32_t], [0_c_int32_t], ptr, 1_c_int32_t)
  end function nc_write_profile

  ! Field `name` over x, y, z with the shape of f (IO_ncf.f90:239-270: the
  ! dimensions must match the grid)
  integer function nc_read_field(fname, name, f) result(rc)
    character(len=*), intent(in) :: fname, name
    real(c_float), intent(out), target, contiguous :: f(:, :, :)
    rc = qg_nc_read(trim(fname)//c_null_char, trim(name)//c_null_char, 'z,y,x'//c_null_char, &
                    [int(size(f, 3), c_int32_t), int(size(f, 2), c_int32_t), int(size(f, 1), c_int32_t)], &
                    3_c_int32_t, c_loc(f))
  end function nc_read_field

  ! The one-element time over t
  integer function nc_read_time(fname, time) result(rc)
    character(len=*), intent(in) :: fname
    real(c_float), intent(out) :: time
    real(c_float), target :: t(1)
    rc = qg_nc_read(trim(fname)//c_null_char, 'time'//c_null_char, 't'//c_null_char, [1_c_int32_t], 1_c_int32_t, &
                    c_loc(t))
    time = t(1)
  end function nc_read_time

end module qg_ncio
