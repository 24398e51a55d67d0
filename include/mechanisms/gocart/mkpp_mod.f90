!==============================================================================
! mkpp_mod.f90 -- Fortran binding of the compiled mechanism gocart (libmkpp_gocart.so).
! Generated by mkpp_amd.gen.emit_host; do not edit.
!==============================================================================
module mkpp_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private

  integer(c_int), parameter, public :: MKPP_SUCCESS = 0_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_NULL_PTR = -101_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_INVALID_HDL = -102_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_INVALID_ARG = -103_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_INVALID_STE = -104_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_CONVERGENCE = -105_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_NOT_FOUND = -106_c_int

  integer, parameter, public :: MKPP_NUM_RATE_PARAMETERS = 1
  integer, parameter, public :: MKPP_NUM_SPECIES = 26
  integer, parameter, public :: MKPP_NUM_REACTIONS = 27
  integer, parameter, public :: MKPP_NUM_PHOTOLYSIS = 1
  ! generation-time solver options of this drop-in (method id 0 ROS-2 ... 4 RODAS-4)
  integer, parameter, public :: MKPP_DEFAULT_METHOD = 1

  type, public :: mkpp_handle_t
    type(c_ptr) :: ptr = c_null_ptr
  end type mkpp_handle_t

  public :: mkpp_create_handle
  public :: mkpp_destroy_handle
  public :: mkpp_initialize
  public :: mkpp_get_species_count
  public :: mkpp_get_reaction_count
  public :: mkpp_get_photolysis_count
  public :: mkpp_get_species_index
  public :: mkpp_get_photolysis_index
  public :: mkpp_set_state_ptrs
  public :: mkpp_set_rate_parameter_ptrs
  public :: mkpp_set_photolysis_ptrs
  public :: mkpp_set_cloud_water_ptr
  public :: mkpp_set_aerosol_ptrs
  public :: mkpp_integrate
  public :: mkpp_evaluate_rates
  public :: mkpp_evaluate_jacobian
  public :: mkpp_get_last_error

  interface
    integer(c_int) function c_create_handle(handle_out, num_cells) bind(C, name="mkpp_create_handle")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), intent(out) :: handle_out
      integer(c_size_t), value :: num_cells
    end function c_create_handle
    integer(c_int) function c_destroy_handle(handle) bind(C, name="mkpp_destroy_handle")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
    end function c_destroy_handle
    integer(c_int) function c_initialize(handle) bind(C, name="mkpp_initialize")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
    end function c_initialize
    integer(c_int) function c_get_species_count(handle, count_out) bind(C, name="mkpp_get_species_count")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      integer(c_size_t), intent(out) :: count_out
    end function c_get_species_count
    integer(c_int) function c_get_reaction_count(handle, count_out) bind(C, name="mkpp_get_reaction_count")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      integer(c_size_t), intent(out) :: count_out
    end function c_get_reaction_count
    integer(c_int) function c_get_photolysis_count(handle, count_out) bind(C, name="mkpp_get_photolysis_count")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      integer(c_size_t), intent(out) :: count_out
    end function c_get_photolysis_count
    integer(c_int) function c_get_species_index(handle, name, index_out) bind(C, name="mkpp_get_species_index")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int), intent(out) :: index_out
    end function c_get_species_index
    integer(c_int) function c_get_photolysis_index(handle, name, index_out) bind(C, name="mkpp_get_photolysis_index")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int), intent(out) :: index_out
    end function c_get_photolysis_index
    integer(c_int) function c_set_state_ptrs(handle, conc, temp, pres, rho) bind(C, name="mkpp_set_state_ptrs")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      type(c_ptr), value :: conc
      type(c_ptr), value :: temp
      type(c_ptr), value :: pres
      type(c_ptr), value :: rho
    end function c_set_state_ptrs
    integer(c_int) function c_set_rate_parameter_ptrs(handle, values, count) bind(C, name="mkpp_set_rate_parameter_ptrs")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      type(c_ptr), value :: values
      integer(c_size_t), value :: count
    end function c_set_rate_parameter_ptrs
    integer(c_int) function c_set_photolysis_ptrs(handle, values) bind(C, name="mkpp_set_photolysis_ptrs")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      type(c_ptr), value :: values
    end function c_set_photolysis_ptrs
    integer(c_int) function c_set_cloud_water_ptr(handle, values) bind(C, name="mkpp_set_cloud_water_ptr")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      type(c_ptr), value :: values
    end function c_set_cloud_water_ptr
    integer(c_int) function c_set_aerosol_ptrs(handle, rh, sa) bind(C, name="mkpp_set_aerosol_ptrs")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      type(c_ptr), value :: rh
      type(c_ptr), value :: sa
    end function c_set_aerosol_ptrs
    integer(c_int) function c_integrate(handle, dt) bind(C, name="mkpp_integrate")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      real(c_double), value :: dt
    end function c_integrate
    integer(c_int) function c_evaluate_rates(handle, out) bind(C, name="mkpp_evaluate_rates")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      type(c_ptr), value :: out
    end function c_evaluate_rates
    integer(c_int) function c_evaluate_jacobian(handle, out) bind(C, name="mkpp_evaluate_jacobian")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      type(c_ptr), value :: out
    end function c_evaluate_jacobian
    integer(c_int) function c_get_last_error(handle, buffer, buffer_len) bind(C, name="mkpp_get_last_error")
      import :: c_ptr, c_int, c_size_t, c_double, c_char
      type(c_ptr), value :: handle
      character(kind=c_char), intent(out) :: buffer(*)
      integer(c_size_t), value :: buffer_len
    end function c_get_last_error
  end interface

contains

  subroutine mkpp_create_handle(handle, num_cells, status)
    type(mkpp_handle_t), intent(out) :: handle
    integer, intent(in) :: num_cells
    integer, intent(out) :: status
    status = int(c_create_handle(handle%ptr, int(num_cells, c_size_t)))
  end subroutine mkpp_create_handle

  subroutine mkpp_destroy_handle(handle, status)
    type(mkpp_handle_t), intent(inout) :: handle
    integer, intent(out) :: status
    status = int(c_destroy_handle(handle%ptr))
    handle%ptr = c_null_ptr
  end subroutine mkpp_destroy_handle

  subroutine mkpp_initialize(handle, status)
    type(mkpp_handle_t), intent(in) :: handle
    integer, intent(out) :: status
    status = int(c_initialize(handle%ptr))
  end subroutine mkpp_initialize

  subroutine mkpp_get_species_count(handle, count_out, status)
    type(mkpp_handle_t), intent(in) :: handle
    integer, intent(out) :: count_out
    integer, intent(out) :: status
    integer(c_size_t) :: n
    status = int(c_get_species_count(handle%ptr, n))
    count_out = int(n)
  end subroutine mkpp_get_species_count

  subroutine mkpp_get_reaction_count(handle, count_out, status)
    type(mkpp_handle_t), intent(in) :: handle
    integer, intent(out) :: count_out
    integer, intent(out) :: status
    integer(c_size_t) :: n
    status = int(c_get_reaction_count(handle%ptr, n))
    count_out = int(n)
  end subroutine mkpp_get_reaction_count

  subroutine mkpp_get_photolysis_count(handle, count_out, status)
    type(mkpp_handle_t), intent(in) :: handle
    integer, intent(out) :: count_out
    integer, intent(out) :: status
    integer(c_size_t) :: n
    status = int(c_get_photolysis_count(handle%ptr, n))
    count_out = int(n)
  end subroutine mkpp_get_photolysis_count

  subroutine mkpp_get_species_index(handle, name, index_out, status)
    type(mkpp_handle_t), intent(in) :: handle
    character(len=*), intent(in) :: name
    integer, intent(out) :: index_out
    integer, intent(out) :: status
    integer(c_int) :: i0
    status = int(c_get_species_index(handle%ptr, trim(name) // c_null_char, i0))
    index_out = -1
    if (status == MKPP_SUCCESS) index_out = int(i0) + 1  ! 1-based
  end subroutine mkpp_get_species_index

  subroutine mkpp_get_photolysis_index(handle, name, index_out, status)
    type(mkpp_handle_t), intent(in) :: handle
    character(len=*), intent(in) :: name
    integer, intent(out) :: index_out
    integer, intent(out) :: status
    integer(c_int) :: i0
    status = int(c_get_photolysis_index(handle%ptr, trim(name) // c_null_char, i0))
    index_out = -1
    if (status == MKPP_SUCCESS) index_out = int(i0) + 1  ! 1-based
  end subroutine mkpp_get_photolysis_index

  subroutine mkpp_set_state_ptrs(handle, conc, temp, pres, rho, status)
    type(mkpp_handle_t), intent(in) :: handle
    real(c_double), target, intent(inout) :: conc(:, :)  ! (cells, species)
    real(c_double), target, intent(in) :: temp(:), pres(:), rho(:)
    integer, intent(out) :: status
    if (.not. (is_contiguous(conc) .and. is_contiguous(temp) .and. is_contiguous(pres) .and. &
               is_contiguous(rho))) then
      status = MKPP_ERROR_INVALID_ARG
      return
    end if
    status = int(c_set_state_ptrs(handle%ptr, c_loc(conc), c_loc(temp), c_loc(pres), c_loc(rho)))
  end subroutine mkpp_set_state_ptrs

  subroutine mkpp_set_rate_parameter_ptrs(handle, values, status)
    type(mkpp_handle_t), intent(in) :: handle
    real(c_double), target, intent(in) :: values(:, :)  ! (cells, MKPP_NUM_RATE_PARAMETERS)
    integer, intent(out) :: status
    if (.not. is_contiguous(values) .or. size(values, 2) /= MKPP_NUM_RATE_PARAMETERS) then
      status = MKPP_ERROR_INVALID_ARG
      return
    end if
    status = int(c_set_rate_parameter_ptrs(handle%ptr, c_loc(values), int(size(values, 2), c_size_t)))
  end subroutine mkpp_set_rate_parameter_ptrs

  subroutine mkpp_set_photolysis_ptrs(handle, values, status)
    type(mkpp_handle_t), intent(in) :: handle
    real(c_double), target, intent(in) :: values(:, :)  ! (cells, MKPP_NUM_PHOTOLYSIS)
    integer, intent(out) :: status
    if (.not. is_contiguous(values)) then
      status = MKPP_ERROR_INVALID_ARG
      return
    end if
    status = int(c_set_photolysis_ptrs(handle%ptr, c_loc(values)))
  end subroutine mkpp_set_photolysis_ptrs

  subroutine mkpp_set_cloud_water_ptr(handle, values, status)
    type(mkpp_handle_t), intent(in) :: handle
    real(c_double), target, intent(in) :: values(:)
    integer, intent(out) :: status
    status = int(c_set_cloud_water_ptr(handle%ptr, c_loc(values)))
  end subroutine mkpp_set_cloud_water_ptr

  subroutine mkpp_set_aerosol_ptrs(handle, rh, sa, status)
    type(mkpp_handle_t), intent(in) :: handle
    real(c_double), target, intent(in) :: rh(:), sa(:)
    integer, intent(out) :: status
    status = int(c_set_aerosol_ptrs(handle%ptr, c_loc(rh), c_loc(sa)))
  end subroutine mkpp_set_aerosol_ptrs

  subroutine mkpp_integrate(handle, dt, status)
    type(mkpp_handle_t), intent(in) :: handle
    real(c_double), intent(in) :: dt
    integer, intent(out) :: status
    status = int(c_integrate(handle%ptr, dt))
  end subroutine mkpp_integrate

  subroutine mkpp_evaluate_rates(handle, rates, status)
    type(mkpp_handle_t), intent(in) :: handle
    real(c_double), target, intent(out) :: rates(:, :)  ! (cells, species)
    integer, intent(out) :: status
    status = int(c_evaluate_rates(handle%ptr, c_loc(rates)))
  end subroutine mkpp_evaluate_rates

  subroutine mkpp_evaluate_jacobian(handle, jac, status)
    type(mkpp_handle_t), intent(in) :: handle
    real(c_double), target, intent(out) :: jac(:, :)  ! (cells, species * species), column j*N + i
    integer, intent(out) :: status
    status = int(c_evaluate_jacobian(handle%ptr, c_loc(jac)))
  end subroutine mkpp_evaluate_jacobian

  subroutine mkpp_get_last_error(handle, message)
    type(mkpp_handle_t), intent(in) :: handle
    character(len=*), intent(out) :: message
    character(kind=c_char) :: buf(512)
    integer(c_int) :: rc
    integer :: i
    buf = c_null_char
    rc = c_get_last_error(handle%ptr, buf, 512_c_size_t)
    message = ''
    do i = 1, min(len(message), 512)
      if (buf(i) == c_null_char) exit
      message(i:i) = buf(i)
    end do
  end subroutine mkpp_get_last_error

end module mkpp_mod
