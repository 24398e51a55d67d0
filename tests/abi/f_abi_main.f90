! A Fortran caller of the drop-in, compiled with flang against the generated
! mkpp_mod.f90 and linked to libmkpp_<mechanism>.so.  Arrays are
! (cells, species) and (cells, parameters), sized with the module's
! MKPP_NUM_* parameters, as a host model declares them.
!
! usage: f_abi_main IN OUT DT
!   IN:  the c_abi_main input format (int64 cells, then float64 conc, temp,
!        pressure, air density and rate parameters, cell index fastest)
!   OUT: int32 return code of mkpp_integrate, int32 1-based index of the
!        species named by the fourth argument (or -1), int32 the status of a
!        set_rate_parameter_ptrs call with the wrong second extent (must be
!        MKPP_ERROR_INVALID_ARG), then the integrated conc
program f_abi_main
  use, intrinsic :: iso_c_binding, only: c_double, c_int32_t, c_int64_t
  use mkpp_mod
  implicit none
  character(len=512) :: path_in, path_out, arg, species, msg
  integer(c_int64_t) :: n64
  integer :: n, status, ns, idx, bad_status, u
  real(c_double) :: dt
  real(c_double), allocatable :: conc(:, :), temp(:), pres(:), rho(:), params(:, :), wrong(:, :)
  type(mkpp_handle_t) :: h

  if (command_argument_count() /= 4) stop 2
  call get_command_argument(1, path_in)
  call get_command_argument(2, path_out)
  call get_command_argument(3, arg)
  call get_command_argument(4, species)
  read (arg, *) dt

  open (newunit=u, file=trim(path_in), access='stream', form='unformatted', status='old', action='read')
  read (u) n64
  n = int(n64)
  allocate (conc(n, MKPP_NUM_SPECIES), temp(n), pres(n), rho(n), params(n, max(MKPP_NUM_RATE_PARAMETERS, 1)))
  allocate (wrong(n, MKPP_NUM_RATE_PARAMETERS + 1))
  read (u) conc, temp, pres, rho
  if (MKPP_NUM_RATE_PARAMETERS > 0) read (u) params(:, 1:MKPP_NUM_RATE_PARAMETERS)
  close (u)

  idx = -1
  bad_status = MKPP_SUCCESS
  call mkpp_create_handle(h, n, status)
  if (status == MKPP_SUCCESS) call mkpp_initialize(h, status)
  if (status == MKPP_SUCCESS) call mkpp_get_species_count(h, ns, status)
  if (status == MKPP_SUCCESS .and. ns /= MKPP_NUM_SPECIES) status = -1
  if (status == MKPP_SUCCESS) call mkpp_get_species_index(h, trim(species), idx, status)
  if (status == MKPP_SUCCESS) call mkpp_set_state_ptrs(h, conc, temp, pres, rho, status)
  if (status == MKPP_SUCCESS) call mkpp_set_rate_parameter_ptrs(h, wrong, bad_status)
  if (status == MKPP_SUCCESS .and. MKPP_NUM_RATE_PARAMETERS > 0) &
    call mkpp_set_rate_parameter_ptrs(h, params(:, 1:MKPP_NUM_RATE_PARAMETERS), status)
  if (status == MKPP_SUCCESS) call mkpp_integrate(h, dt, status)
  if (status /= MKPP_SUCCESS) then
    call mkpp_get_last_error(h, msg)
    print '(a, i0, 2a)', 'mkpp error ', status, ': ', trim(msg)
  end if
  call mkpp_destroy_handle(h, ns)

  open (newunit=u, file=trim(path_out), access='stream', form='unformatted', status='replace', action='write')
  write (u) int(status, c_int32_t), int(idx, c_int32_t), int(bad_status, c_int32_t)
  write (u) conc
  close (u)
  print '(a, i0, a, i0, a, i0, a, i0)', 'f_abi_main: ', n, ' cells of ', MKPP_NUM_SPECIES, &
    ' species, method ', MKPP_DEFAULT_METHOD, ', status ', status
  if (status /= MKPP_SUCCESS) stop 1
end program f_abi_main
