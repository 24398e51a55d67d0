! A Fortran caller of the device API through include/mkpp_amd_mod.f90,
! compiled with flang and linked to libmkpp_amd.so.  Device memory comes from
! HIP's C API (bound here as a host model binds it, e.g. through hipfort).
!
! usage: f_device_main IN OUT
!   IN:  the c_abi_main input format (int64 cells, then float64 conc, temp,
!        pressure, air density and rate parameters, cell index fastest)
!   OUT: four runs of saprc99 over 3600 s with Rodas3, atol 1e3, rtol 1e-3,
!        seed 1e-6, each int32 return code, the integrated conc and int32
!        status per cell:
!     1. mkpp_amd_device_integrate on a stream of its own;
!     2. mkpp_amd_device_integrate_batch_static, the cells split into two
!        segments (sub-arrays of the same device buffers);
!     3. mkpp_amd_device_integrate_batch with a workspace, the same segments;
!     4. mkpp_amd_execute on a handle with those solver options, per-cell
!        temperature, density and forcing (status from get_cell_status).
program f_device_main
  use, intrinsic :: iso_c_binding
  use mkpp_amd_mod
  implicit none

  interface
    integer(c_int) function hipMalloc(p, bytes) bind(C, name="hipMalloc")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), intent(out) :: p
      integer(c_size_t), value :: bytes
    end function
    integer(c_int) function hipFree(p) bind(C, name="hipFree")
      import :: c_int, c_ptr
      type(c_ptr), value :: p
    end function
    ! kind: 1 host to device, 2 device to host
    integer(c_int) function hipMemcpy(dst, src, bytes, kind) bind(C, name="hipMemcpy")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: dst, src
      integer(c_size_t), value :: bytes
      integer(c_int), value :: kind
    end function
    integer(c_int) function hipStreamCreate(s) bind(C, name="hipStreamCreate")
      import :: c_int, c_ptr
      type(c_ptr), intent(out) :: s
    end function
    integer(c_int) function hipStreamSynchronize(s) bind(C, name="hipStreamSynchronize")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function hipStreamDestroy(s) bind(C, name="hipStreamDestroy")
      import :: c_int, c_ptr
      type(c_ptr), value :: s
    end function
    integer(c_int) function hipDeviceSynchronize() bind(C, name="hipDeviceSynchronize")
      import :: c_int
    end function
  end interface

  character(len=512) :: path_in, path_out, msg
  integer(c_int64_t) :: n64
  integer(c_long) :: n, half
  integer(c_int) :: ns, nr, nph, npar, env, rc
  integer :: u, run, status
  real(c_double), allocatable, target :: conc0(:, :), conc(:, :), temp(:), pres(:), rho(:), params(:, :)
  real(c_double), allocatable, target :: atol(:), rtol(:)
  integer(c_int), allocatable, target :: cell_status(:)
  type(c_ptr) :: d_conc, d_temp, d_rho, d_params, d_atol, d_rtol, d_status, d_work, stream, handle
  type(mkpp_amd_batch_segment) :: seg(2)
  type(mkpp_amd_execution_params) :: ep
  integer(c_size_t) :: cbytes, nbytes
  integer(c_intptr_t), parameter :: F64 = 8, I32 = 4

  if (command_argument_count() /= 2) stop 2
  call get_command_argument(1, path_in)
  call get_command_argument(2, path_out)
  rc = mkpp_amd_mechanism_info("saprc99" // c_null_char, ns, nr, nph, npar, env)
  if (rc /= MKPP_SUCCESS) stop 3

  open (newunit=u, file=trim(path_in), access='stream', form='unformatted', status='old', action='read')
  read (u) n64
  n = int(n64, c_long)
  allocate (conc0(n, ns), conc(n, ns), temp(n), pres(n), rho(n), params(n, npar), cell_status(n))
  allocate (atol(ns), rtol(ns))
  read (u) conc0, temp, pres, rho, params
  close (u)
  atol = 1.0e3_c_double
  rtol = 1.0e-3_c_double
  half = n / 2

  cbytes = int(8 * n * ns, c_size_t)
  nbytes = int(8 * n, c_size_t)
  call check(hipMalloc(d_conc, cbytes))
  call check(hipMalloc(d_temp, nbytes))
  call check(hipMalloc(d_rho, nbytes))
  call check(hipMalloc(d_params, int(8 * n * npar, c_size_t)))
  call check(hipMalloc(d_atol, int(8 * ns, c_size_t)))
  call check(hipMalloc(d_rtol, int(8 * ns, c_size_t)))
  call check(hipMalloc(d_status, int(4 * n, c_size_t)))
  call check(hipMalloc(d_work, MKPP_AMD_BATCH_WORKSPACE_BYTES))
  call check(hipMemcpy(d_temp, c_loc(temp), nbytes, 1))
  call check(hipMemcpy(d_rho, c_loc(rho), nbytes, 1))
  call check(hipMemcpy(d_params, c_loc(params), int(8 * n * npar, c_size_t), 1))
  call check(hipMemcpy(d_atol, c_loc(atol), int(8 * ns, c_size_t), 1))
  call check(hipMemcpy(d_rtol, c_loc(rtol), int(8 * ns, c_size_t), 1))
  call check(hipStreamCreate(stream))

  ! two segments over cells [0, half) and [half, n) of the same buffers
  seg(1)%mechanism = mkpp_amd_mechanism_cptr("saprc99")
  if (.not. c_associated(seg(1)%mechanism)) stop 4
  seg(1)%conc = d_conc
  seg(1)%num_cells = half
  seg(1)%stride = n
  seg(1)%temp = d_temp
  seg(1)%air_density = d_rho
  seg(1)%rate_parameters = d_params
  seg(1)%parameter_stride = n
  seg(1)%atol = d_atol
  seg(1)%rtol = d_rtol
  seg(1)%status = d_status
  seg(2) = seg(1)
  seg(2)%conc = mkpp_amd_offset(d_conc, F64 * half)
  seg(2)%num_cells = n - half
  seg(2)%temp = mkpp_amd_offset(d_temp, F64 * half)
  seg(2)%air_density = mkpp_amd_offset(d_rho, F64 * half)
  seg(2)%rate_parameters = mkpp_amd_offset(d_params, F64 * half)
  seg(2)%status = mkpp_amd_offset(d_status, I32 * half)

  open (newunit=u, file=trim(path_out), access='stream', form='unformatted', status='replace', action='write')
  do run = 1, 4
    call check(hipMemcpy(d_conc, c_loc(conc0), cbytes, 1))
    select case (run)
    case (1)
      rc = mkpp_amd_device_integrate("saprc99" // c_null_char, d_conc, n, n, d_temp, d_rho, d_params, n, &
                                     3600.0_c_double, 1_c_int, MKPP_METHOD_RODAS3, 1.0e-6_c_double, d_atol, d_rtol, &
                                     d_status, c_null_ptr, c_null_ptr, 100000_c_long, stream)
      call check(hipStreamSynchronize(stream))
    case (2)
      rc = mkpp_amd_device_integrate_batch_static(seg, 2_c_int, 3600.0_c_double, 1_c_int, MKPP_METHOD_RODAS3, &
                                                  1.0e-6_c_double, 100000_c_long, 1_c_int, 0_c_int, stream)
      call check(hipStreamSynchronize(stream))
    case (3)
      rc = mkpp_amd_device_integrate_batch(seg, 2_c_int, 3600.0_c_double, 1_c_int, MKPP_METHOD_RODAS3, &
                                           1.0e-6_c_double, 100000_c_long, 1_c_int, 0_c_int, d_work, stream)
      call check(hipStreamSynchronize(stream))
    case (4)
      call mkpp_amd_create_handle_f("saprc99", int(n), handle, status)
      if (status /= MKPP_SUCCESS) stop 5
      rc = mkpp_amd_initialize(handle)
      if (rc == MKPP_SUCCESS) &
        rc = mkpp_amd_set_solver_options(handle, MKPP_METHOD_RODAS3, c_loc(atol), c_loc(rtol), 1.0e-6_c_double)
      if (rc == MKPP_SUCCESS) then
        ep%dt = 3600.0_c_double
        ep%cell_temperature = d_temp
        ep%cell_density = d_rho
        ep%cell_jvals = d_params
        ep%cell_jvals_extent0 = n
        ep%cell_jvals_extent1 = npar
        ep%cell_jvals_stride0 = 1
        ep%cell_jvals_stride1 = n
        rc = mkpp_amd_execute(handle, d_conc, n, n, ep, stream)
      end if
      if (rc /= MKPP_SUCCESS) then
        call mkpp_amd_last_error_f(handle, msg)
        print '(a, i0, 2a)', 'execute: ', rc, ' ', trim(msg)
      end if
    end select
    call check(hipDeviceSynchronize())
    call check(hipMemcpy(c_loc(conc), d_conc, cbytes, 2))
    if (run == 4) then
      if (mkpp_amd_get_cell_status(handle, c_loc(cell_status), c_null_ptr, c_null_ptr) /= MKPP_SUCCESS) stop 6
      if (mkpp_amd_destroy_handle(handle) /= MKPP_SUCCESS) stop 8
    else
      call check(hipMemcpy(c_loc(cell_status), d_status, int(4 * n, c_size_t), 2))
    end if
    write (u) int(rc, c_int32_t), conc, cell_status
    print '(a, i0, a, i0)', 'f_device_main: run ', run, ' rc ', rc
  end do
  close (u)
  call check(hipStreamDestroy(stream))
  call check(hipFree(d_conc))
  call check(hipFree(d_temp))
  call check(hipFree(d_rho))
  call check(hipFree(d_params))
  call check(hipFree(d_atol))
  call check(hipFree(d_rtol))
  call check(hipFree(d_status))
  call check(hipFree(d_work))
  print '(a, i0, a)', 'f_device_main: ', n, ' cells done'

contains

  subroutine check(code)
    integer(c_int), intent(in) :: code
    if (code /= 0) then
      print '(a, i0)', 'HIP call failed: ', code
      stop 7
    end if
  end subroutine check

end program f_device_main
