!==============================================================================
! mkpp_amd_mod.f90 -- Fortran binding of libmkpp_amd.so (mkpp_amd.h).
!
! The reference's Fortran module binds the per-mechanism host C ABI with
! bind(C) interfaces and status-returning wrappers
! (src/mkpp/templates/host_api/mkpp_mod.f90.j2:54-169); the generated
! include/mechanisms/<mech>/mkpp_mod.f90 does the same for each drop-in.
! This module binds the entry points that have no reference counterpart:
!
!   * the registry (mechanism names, metadata);
!   * the handle API with the mechanism chosen at creation, and its solver
!     option, step and per-cell status extensions;
!   * the stream-ordered device API on device pointers (integrate, reduced
!     integrate, the unrolled validation backend, rates/Jacobian), the persistent multi-mechanism batch (with
!     and without a workspace), registry-style execute, and the
!     differentiation entry points (checkpointed forward, adjoint,
!     tangent-linear).
!
! Device pointers and HIP streams are type(c_ptr) (as hipfort passes them);
! c_null_ptr selects the default stream or an absent optional array.  Index
! arguments keep the C ABI's 0-based convention except in the wrappers named
! *_f, which take and return Fortran strings and 1-based indices.  Arrays on
! the device are laid out as the C ABI states: conc(cell + stride*species),
! i.e. a Fortran (stride, species) array.
!==============================================================================
module mkpp_amd_mod
  use, intrinsic :: iso_c_binding
  implicit none
  private

  ! return codes (mkpp.h)
  integer(c_int), parameter, public :: MKPP_SUCCESS = 0_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_NULL_PTR = -101_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_INVALID_HDL = -102_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_INVALID_ARG = -103_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_INVALID_STE = -104_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_CONVERGENCE = -105_c_int
  integer(c_int), parameter, public :: MKPP_ERROR_NOT_FOUND = -106_c_int

  ! method ids
  integer(c_int), parameter, public :: MKPP_METHOD_ROS2 = 0_c_int
  integer(c_int), parameter, public :: MKPP_METHOD_ROS3 = 1_c_int
  integer(c_int), parameter, public :: MKPP_METHOD_ROS4 = 2_c_int
  integer(c_int), parameter, public :: MKPP_METHOD_RODAS3 = 3_c_int
  integer(c_int), parameter, public :: MKPP_METHOD_RODAS4 = 4_c_int

  ! per-cell status codes
  integer(c_int), parameter, public :: MKPP_CELL_SUCCESS = 0_c_int
  integer(c_int), parameter, public :: MKPP_CELL_ERR_SINGULAR_MATRIX = 1_c_int
  integer(c_int), parameter, public :: MKPP_CELL_ERR_NON_FINITE_STATE = 2_c_int
  integer(c_int), parameter, public :: MKPP_CELL_ERR_CONVERGENCE_FAILURE = 3_c_int
  integer(c_int), parameter, public :: MKPP_CELL_ERR_NUMERICAL_FAILURE = 4_c_int
  integer(c_int), parameter, public :: MKPP_CELL_ERR_CHECKPOINT_OVERFLOW = 5_c_int

  ! batch limits
  integer(c_int), parameter, public :: MKPP_AMD_BATCH_MAX_SEGMENTS = 8_c_int
  integer(c_size_t), parameter, public :: MKPP_AMD_BATCH_WORKSPACE_BYTES = 32_c_size_t
  integer(c_size_t), parameter, public :: MKPP_AMD_BATCH_STATIC_SCRATCH_BYTES = 256_c_size_t

  ! one segment of a persistent batch (mkpp_amd_batch_segment); `mechanism`
  ! must point at a NUL-terminated name that outlives the call:
  ! mkpp_amd_mechanism_cptr returns the library's own copy
  type, bind(C), public :: mkpp_amd_batch_segment
    type(c_ptr) :: mechanism = c_null_ptr
    type(c_ptr) :: conc = c_null_ptr
    integer(c_long) :: num_cells = 0
    integer(c_long) :: stride = 0
    type(c_ptr) :: temp = c_null_ptr
    type(c_ptr) :: air_density = c_null_ptr
    type(c_ptr) :: rate_parameters = c_null_ptr
    integer(c_long) :: parameter_stride = 0
    type(c_ptr) :: atol = c_null_ptr
    type(c_ptr) :: rtol = c_null_ptr
    type(c_ptr) :: status = c_null_ptr
    type(c_ptr) :: accepted = c_null_ptr
    type(c_ptr) :: attempted = c_null_ptr
  end type mkpp_amd_batch_segment

  ! registry-style execution parameters (mkpp_amd_execution_params); the
  ! defaults are those of the reference's HostExecutionParams
  type, bind(C), public :: mkpp_amd_execution_params
    real(c_double) :: dt = 60.0_c_double
    integer(c_int) :: steps = 1_c_int
    type(c_ptr) :: jvals = c_null_ptr
    integer(c_long) :: jvals_extent = 0
    real(c_double) :: temp = 288.15_c_double
    real(c_double) :: air_density = 2.4476e19_c_double
    real(c_double) :: rh = 0.5_c_double
    real(c_double) :: clw = 0.0_c_double
    type(c_ptr) :: cell_temperature = c_null_ptr
    type(c_ptr) :: cell_humidity = c_null_ptr
    type(c_ptr) :: cell_density = c_null_ptr
    type(c_ptr) :: cell_cloud_water = c_null_ptr
    type(c_ptr) :: cell_jvals = c_null_ptr
    integer(c_long) :: cell_jvals_extent0 = 0
    integer(c_long) :: cell_jvals_extent1 = 0
    integer(c_long) :: cell_jvals_stride0 = 0
    integer(c_long) :: cell_jvals_stride1 = 0
  end type mkpp_amd_execution_params

  public :: mkpp_amd_mechanism_count, mkpp_amd_mechanism_name, mkpp_amd_has_mechanism, mkpp_amd_mechanism_info
  public :: mkpp_amd_species_name
  public :: mkpp_amd_create_handle, mkpp_amd_destroy_handle, mkpp_amd_initialize, mkpp_amd_get_species_count
  public :: mkpp_amd_get_reaction_count, mkpp_amd_get_photolysis_count, mkpp_amd_get_species_index
  public :: mkpp_amd_get_photolysis_index, mkpp_amd_set_state_ptrs, mkpp_amd_set_rate_parameter_ptrs
  public :: mkpp_amd_set_photolysis_ptrs, mkpp_amd_set_cloud_water_ptr, mkpp_amd_set_aerosol_ptrs
  public :: mkpp_amd_integrate, mkpp_amd_evaluate_rates, mkpp_amd_evaluate_jacobian, mkpp_amd_get_last_error
  public :: mkpp_amd_set_solver_options, mkpp_amd_get_solver_options, mkpp_amd_set_max_attempts
  public :: mkpp_amd_integrate_steps, mkpp_amd_get_cell_status
  public :: mkpp_amd_device_integrate, mkpp_amd_device_integrate_reduced, mkpp_amd_device_evaluate
  public :: mkpp_amd_device_integrate_unrolled
  public :: mkpp_amd_device_integrate_batch, mkpp_amd_device_integrate_batch_static, mkpp_amd_batch_capacity
  public :: mkpp_amd_execute
  public :: mkpp_amd_diff_info, mkpp_amd_device_rate_constants, mkpp_amd_device_integrate_checkpointed
  public :: mkpp_amd_device_adjoint, mkpp_amd_device_tangent_linear
  public :: mkpp_amd_c_string, mkpp_amd_mechanism_cptr, mkpp_amd_create_handle_f, mkpp_amd_species_index_f
  public :: mkpp_amd_last_error_f, mkpp_amd_offset

  interface
    ! ---- registry ----
    integer(c_int) function mkpp_amd_mechanism_count() bind(C, name="mkpp_amd_mechanism_count")
      import :: c_int
    end function
    type(c_ptr) function mkpp_amd_mechanism_name(index) bind(C, name="mkpp_amd_mechanism_name")
      import :: c_ptr, c_int
      integer(c_int), value :: index
    end function
    integer(c_int) function mkpp_amd_has_mechanism(name) bind(C, name="mkpp_amd_has_mechanism")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: name(*)
    end function
    integer(c_int) function mkpp_amd_mechanism_info(name, num_species, num_reactions, num_photolysis, &
                                                    num_rate_parameters, supports_runtime_environment) &
        bind(C, name="mkpp_amd_mechanism_info")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int), intent(out) :: num_species, num_reactions, num_photolysis, num_rate_parameters
      integer(c_int), intent(out) :: supports_runtime_environment
    end function
    integer(c_int) function mkpp_amd_species_name(mechanism, index, name_out) bind(C, name="mkpp_amd_species_name")
      import :: c_int, c_char, c_ptr
      character(kind=c_char), intent(in) :: mechanism(*)
      integer(c_int), value :: index
      type(c_ptr), intent(out) :: name_out
    end function

    ! ---- handle API (host pointers, mkpp.h meaning) ----
    integer(c_int) function mkpp_amd_create_handle(mechanism, handle_out, num_cells) &
        bind(C, name="mkpp_amd_create_handle")
      import :: c_int, c_char, c_ptr, c_size_t
      character(kind=c_char), intent(in) :: mechanism(*)
      type(c_ptr), intent(out) :: handle_out
      integer(c_size_t), value :: num_cells
    end function
    integer(c_int) function mkpp_amd_destroy_handle(handle) bind(C, name="mkpp_amd_destroy_handle")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function
    integer(c_int) function mkpp_amd_initialize(handle) bind(C, name="mkpp_amd_initialize")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle
    end function
    integer(c_int) function mkpp_amd_get_species_count(handle, count_out) bind(C, name="mkpp_amd_get_species_count")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: handle
      integer(c_size_t), intent(out) :: count_out
    end function
    integer(c_int) function mkpp_amd_get_reaction_count(handle, count_out) &
        bind(C, name="mkpp_amd_get_reaction_count")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: handle
      integer(c_size_t), intent(out) :: count_out
    end function
    integer(c_int) function mkpp_amd_get_photolysis_count(handle, count_out) &
        bind(C, name="mkpp_amd_get_photolysis_count")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: handle
      integer(c_size_t), intent(out) :: count_out
    end function
    integer(c_int) function mkpp_amd_get_species_index(handle, name, index_out) &
        bind(C, name="mkpp_amd_get_species_index")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: handle
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int), intent(out) :: index_out
    end function
    integer(c_int) function mkpp_amd_get_photolysis_index(handle, name, index_out) &
        bind(C, name="mkpp_amd_get_photolysis_index")
      import :: c_int, c_ptr, c_char
      type(c_ptr), value :: handle
      character(kind=c_char), intent(in) :: name(*)
      integer(c_int), intent(out) :: index_out
    end function
    integer(c_int) function mkpp_amd_set_state_ptrs(handle, conc, temp, pressure, air_density) &
        bind(C, name="mkpp_amd_set_state_ptrs")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, conc, temp, pressure, air_density
    end function
    integer(c_int) function mkpp_amd_set_rate_parameter_ptrs(handle, values, count) &
        bind(C, name="mkpp_amd_set_rate_parameter_ptrs")
      import :: c_int, c_ptr, c_size_t
      type(c_ptr), value :: handle, values
      integer(c_size_t), value :: count
    end function
    integer(c_int) function mkpp_amd_set_photolysis_ptrs(handle, values) bind(C, name="mkpp_amd_set_photolysis_ptrs")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, values
    end function
    integer(c_int) function mkpp_amd_set_cloud_water_ptr(handle, values) bind(C, name="mkpp_amd_set_cloud_water_ptr")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, values
    end function
    integer(c_int) function mkpp_amd_set_aerosol_ptrs(handle, rh, sa) bind(C, name="mkpp_amd_set_aerosol_ptrs")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, rh, sa
    end function
    integer(c_int) function mkpp_amd_integrate(handle, dt) bind(C, name="mkpp_amd_integrate")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), value :: dt
    end function
    integer(c_int) function mkpp_amd_evaluate_rates(handle, out) bind(C, name="mkpp_amd_evaluate_rates")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, out
    end function
    integer(c_int) function mkpp_amd_evaluate_jacobian(handle, out) bind(C, name="mkpp_amd_evaluate_jacobian")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, out
    end function
    integer(c_int) function mkpp_amd_get_last_error(handle, buffer, buffer_len) &
        bind(C, name="mkpp_amd_get_last_error")
      import :: c_int, c_ptr, c_char, c_size_t
      type(c_ptr), value :: handle
      character(kind=c_char), intent(out) :: buffer(*)
      integer(c_size_t), value :: buffer_len
    end function
    ! atol/rtol: c_null_ptr keeps the current values, else c_loc of
    ! num_species host values in physical order
    integer(c_int) function mkpp_amd_set_solver_options(handle, method, atol, rtol, initial_step_seed) &
        bind(C, name="mkpp_amd_set_solver_options")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      integer(c_int), value :: method
      type(c_ptr), value :: atol, rtol
      real(c_double), value :: initial_step_seed
    end function
    integer(c_int) function mkpp_amd_get_solver_options(handle, method, atol, rtol, initial_step_seed) &
        bind(C, name="mkpp_amd_get_solver_options")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, method, atol, rtol, initial_step_seed
    end function
    integer(c_int) function mkpp_amd_set_max_attempts(handle, max_attempts) bind(C, name="mkpp_amd_set_max_attempts")
      import :: c_int, c_ptr, c_long
      type(c_ptr), value :: handle
      integer(c_long), value :: max_attempts
    end function
    integer(c_int) function mkpp_amd_integrate_steps(handle, dt, steps) bind(C, name="mkpp_amd_integrate_steps")
      import :: c_int, c_ptr, c_double
      type(c_ptr), value :: handle
      real(c_double), value :: dt
      integer(c_int), value :: steps
    end function
    ! host arrays of num_cells, or c_null_ptr
    integer(c_int) function mkpp_amd_get_cell_status(handle, status, accepted, attempted) &
        bind(C, name="mkpp_amd_get_cell_status")
      import :: c_int, c_ptr
      type(c_ptr), value :: handle, status, accepted, attempted
    end function

    ! ---- device API: device pointers, enqueued on `stream` ----
    integer(c_int) function mkpp_amd_device_integrate(mechanism, conc, num_cells, stride, temp, air_density, &
                                                      rate_parameters, parameter_stride, dt, steps, method, &
                                                      initial_step_seed, atol, rtol, status, accepted, attempted, &
                                                      max_attempts, stream) bind(C, name="mkpp_amd_device_integrate")
      import :: c_int, c_ptr, c_char, c_long, c_double
      character(kind=c_char), intent(in) :: mechanism(*)
      type(c_ptr), value :: conc
      integer(c_long), value :: num_cells, stride
      type(c_ptr), value :: temp, air_density, rate_parameters
      integer(c_long), value :: parameter_stride
      real(c_double), value :: dt
      integer(c_int), value :: steps, method
      real(c_double), value :: initial_step_seed
      type(c_ptr), value :: atol, rtol, status, accepted, attempted
      integer(c_long), value :: max_attempts
      type(c_ptr), value :: stream
    end function
    integer(c_int) function mkpp_amd_device_integrate_reduced(mechanism, conc, num_cells, stride, temp, &
                                                              air_density, rate_parameters, parameter_stride, dt, &
                                                              steps, method, importance_threshold, atol, rtol, &
                                                              status, accepted, attempted, max_attempts, stream) &
        bind(C, name="mkpp_amd_device_integrate_reduced")
      import :: c_int, c_ptr, c_char, c_long, c_double
      character(kind=c_char), intent(in) :: mechanism(*)
      type(c_ptr), value :: conc
      integer(c_long), value :: num_cells, stride
      type(c_ptr), value :: temp, air_density, rate_parameters
      integer(c_long), value :: parameter_stride
      real(c_double), value :: dt
      integer(c_int), value :: steps, method
      real(c_double), value :: importance_threshold
      type(c_ptr), value :: atol, rtol, status, accepted, attempted
      integer(c_long), value :: max_attempts
      type(c_ptr), value :: stream
    end function
    ! validation backend over the dense unrolled LU (lane-per-cell mechanisms)
    integer(c_int) function mkpp_amd_device_integrate_unrolled(mechanism, conc, num_cells, stride, temp, &
                                                               air_density, rate_parameters, parameter_stride, dt, &
                                                               steps, method, initial_step_seed, atol, rtol, &
                                                               status, accepted, attempted, max_attempts, stream) &
        bind(C, name="mkpp_amd_device_integrate_unrolled")
      import :: c_int, c_ptr, c_char, c_long, c_double
      character(kind=c_char), intent(in) :: mechanism(*)
      type(c_ptr), value :: conc
      integer(c_long), value :: num_cells, stride
      type(c_ptr), value :: temp, air_density, rate_parameters
      integer(c_long), value :: parameter_stride
      real(c_double), value :: dt
      integer(c_int), value :: steps, method
      real(c_double), value :: initial_step_seed
      type(c_ptr), value :: atol, rtol, status, accepted, attempted
      integer(c_long), value :: max_attempts
      type(c_ptr), value :: stream
    end function
    ! what: 0 rates, 1 dense Jacobian
    integer(c_int) function mkpp_amd_device_evaluate(mechanism, what, conc, num_cells, stride, temp, air_density, &
                                                     rate_parameters, parameter_stride, out, out_stride, stream) &
        bind(C, name="mkpp_amd_device_evaluate")
      import :: c_int, c_ptr, c_char, c_long
      character(kind=c_char), intent(in) :: mechanism(*)
      integer(c_int), value :: what
      type(c_ptr), value :: conc
      integer(c_long), value :: num_cells, stride
      type(c_ptr), value :: temp, air_density, rate_parameters
      integer(c_long), value :: parameter_stride
      type(c_ptr), value :: out
      integer(c_long), value :: out_stride
      type(c_ptr), value :: stream
    end function
    integer(c_int) function mkpp_amd_device_integrate_batch(segments, num_segments, dt, steps, method, &
                                                            initial_step_seed, max_attempts, cells_per_lane, &
                                                            workgroups, workspace, stream) &
        bind(C, name="mkpp_amd_device_integrate_batch")
      import :: c_int, c_ptr, c_long, c_double, mkpp_amd_batch_segment
      type(mkpp_amd_batch_segment), intent(in) :: segments(*)
      integer(c_int), value :: num_segments
      real(c_double), value :: dt
      integer(c_int), value :: steps, method
      real(c_double), value :: initial_step_seed
      integer(c_long), value :: max_attempts
      integer(c_int), value :: cells_per_lane, workgroups
      type(c_ptr), value :: workspace, stream
    end function
    integer(c_int) function mkpp_amd_device_integrate_batch_static(segments, num_segments, dt, steps, method, &
                                                                   initial_step_seed, max_attempts, cells_per_lane, &
                                                                   workgroups, stream) &
        bind(C, name="mkpp_amd_device_integrate_batch_static")
      import :: c_int, c_ptr, c_long, c_double, mkpp_amd_batch_segment
      type(mkpp_amd_batch_segment), intent(in) :: segments(*)
      integer(c_int), value :: num_segments
      real(c_double), value :: dt
      integer(c_int), value :: steps, method
      real(c_double), value :: initial_step_seed
      integer(c_long), value :: max_attempts
      integer(c_int), value :: cells_per_lane, workgroups
      type(c_ptr), value :: stream
    end function
    integer(c_int) function mkpp_amd_batch_capacity(method, per_cu, num_cus) bind(C, name="mkpp_amd_batch_capacity")
      import :: c_int
      integer(c_int), value :: method
      integer(c_int), intent(out) :: per_cu, num_cus
    end function
    ! synchronous; per-cell results through mkpp_amd_get_cell_status
    integer(c_int) function mkpp_amd_execute(handle, conc, num_cells, stride, params, stream) &
        bind(C, name="mkpp_amd_execute")
      import :: c_int, c_ptr, c_long, mkpp_amd_execution_params
      type(c_ptr), value :: handle, conc
      integer(c_long), value :: num_cells, stride
      type(mkpp_amd_execution_params), intent(in) :: params
      type(c_ptr), value :: stream
    end function

    ! ---- differentiation (every mechanism and method) ----
    integer(c_int) function mkpp_amd_diff_info(mechanism, num_groups, num_variable) bind(C, name="mkpp_amd_diff_info")
      import :: c_int, c_char
      character(kind=c_char), intent(in) :: mechanism(*)
      integer(c_int), intent(out) :: num_groups, num_variable
    end function
    integer(c_int) function mkpp_amd_device_rate_constants(mechanism, conc, num_cells, stride, temp, air_density, &
                                                           rate_parameters, parameter_stride, k, k_stride, stream) &
        bind(C, name="mkpp_amd_device_rate_constants")
      import :: c_int, c_ptr, c_char, c_long
      character(kind=c_char), intent(in) :: mechanism(*)
      type(c_ptr), value :: conc
      integer(c_long), value :: num_cells, stride
      type(c_ptr), value :: temp, air_density, rate_parameters
      integer(c_long), value :: parameter_stride
      type(c_ptr), value :: k
      integer(c_long), value :: k_stride
      type(c_ptr), value :: stream
    end function
    integer(c_int) function mkpp_amd_device_integrate_checkpointed(mechanism, conc, num_cells, stride, temp, &
                                                                   air_density, rate_parameters, parameter_stride, k, &
                                                                   k_stride, dt, method, initial_step_seed, atol, &
                                                                   rtol, status, accepted, attempted, max_attempts, &
                                                                   ckpt, hs, nsteps, max_steps, replay, stream) &
        bind(C, name="mkpp_amd_device_integrate_checkpointed")
      import :: c_int, c_ptr, c_char, c_long, c_double
      character(kind=c_char), intent(in) :: mechanism(*)
      type(c_ptr), value :: conc
      integer(c_long), value :: num_cells, stride
      type(c_ptr), value :: temp, air_density, rate_parameters
      integer(c_long), value :: parameter_stride
      type(c_ptr), value :: k
      integer(c_long), value :: k_stride
      real(c_double), value :: dt
      integer(c_int), value :: method
      real(c_double), value :: initial_step_seed
      type(c_ptr), value :: atol, rtol, status, accepted, attempted
      integer(c_long), value :: max_attempts
      type(c_ptr), value :: ckpt, hs, nsteps
      integer(c_int), value :: max_steps, replay
      type(c_ptr), value :: stream
    end function
    integer(c_int) function mkpp_amd_device_adjoint(mechanism, conc, num_cells, stride, temp, air_density, &
                                                    rate_parameters, parameter_stride, k, k_stride, method, ckpt, hs, &
                                                    nsteps, max_steps, lambda, kbar, stream) &
        bind(C, name="mkpp_amd_device_adjoint")
      import :: c_int, c_ptr, c_char, c_long
      character(kind=c_char), intent(in) :: mechanism(*)
      type(c_ptr), value :: conc
      integer(c_long), value :: num_cells, stride
      type(c_ptr), value :: temp, air_density, rate_parameters
      integer(c_long), value :: parameter_stride
      type(c_ptr), value :: k
      integer(c_long), value :: k_stride
      integer(c_int), value :: method
      type(c_ptr), value :: ckpt, hs, nsteps
      integer(c_int), value :: max_steps
      type(c_ptr), value :: lambda, kbar, stream
    end function
    integer(c_int) function mkpp_amd_device_tangent_linear(mechanism, conc, num_cells, stride, temp, air_density, &
                                                           rate_parameters, parameter_stride, k, k_stride, method, hs, &
                                                           nsteps, max_steps, dy, dk, status, stream) &
        bind(C, name="mkpp_amd_device_tangent_linear")
      import :: c_int, c_ptr, c_char, c_long
      character(kind=c_char), intent(in) :: mechanism(*)
      type(c_ptr), value :: conc
      integer(c_long), value :: num_cells, stride
      type(c_ptr), value :: temp, air_density, rate_parameters
      integer(c_long), value :: parameter_stride
      type(c_ptr), value :: k
      integer(c_long), value :: k_stride
      integer(c_int), value :: method
      type(c_ptr), value :: hs, nsteps
      integer(c_int), value :: max_steps
      type(c_ptr), value :: dy, dk, status, stream
    end function

    function c_strlen(s) bind(C, name="strlen")
      import :: c_ptr, c_size_t
      type(c_ptr), value :: s
      integer(c_size_t) :: c_strlen
    end function
  end interface

contains

  ! The Fortran string held by a NUL-terminated C string ('' for c_null_ptr).
  function mkpp_amd_c_string(p) result(s)
    type(c_ptr), intent(in) :: p
    character(len=:), allocatable :: s
    character(kind=c_char), pointer :: chars(:)
    integer :: i, n
    if (.not. c_associated(p)) then
      s = ''
      return
    end if
    n = int(c_strlen(p))
    call c_f_pointer(p, chars, [n])
    allocate (character(len=n) :: s)
    do i = 1, n
      s(i:i) = chars(i)
    end do
  end function mkpp_amd_c_string

  ! The library's own NUL-terminated copy of a mechanism name (valid for the
  ! life of the process), as mkpp_amd_batch_segment%mechanism takes it;
  ! c_null_ptr when the library has no such mechanism.
  function mkpp_amd_mechanism_cptr(name) result(p)
    character(len=*), intent(in) :: name
    type(c_ptr) :: p
    integer(c_int) :: i
    do i = 0, mkpp_amd_mechanism_count() - 1
      p = mkpp_amd_mechanism_name(i)
      if (mkpp_amd_c_string(p) == trim(name)) return
    end do
    p = c_null_ptr
  end function mkpp_amd_mechanism_cptr

  ! A handle for `mechanism` over num_cells cells; status as the C call returns it.
  subroutine mkpp_amd_create_handle_f(mechanism, num_cells, handle, status)
    character(len=*), intent(in) :: mechanism
    integer, intent(in) :: num_cells
    type(c_ptr), intent(out) :: handle
    integer, intent(out) :: status
    handle = c_null_ptr
    status = int(mkpp_amd_create_handle(trim(mechanism) // c_null_char, handle, int(num_cells, c_size_t)))
  end subroutine mkpp_amd_create_handle_f

  ! 1-based species index of `name` (-1 when status /= MKPP_SUCCESS).
  subroutine mkpp_amd_species_index_f(handle, name, index_out, status)
    type(c_ptr), intent(in) :: handle
    character(len=*), intent(in) :: name
    integer, intent(out) :: index_out, status
    integer(c_int) :: i0
    status = int(mkpp_amd_get_species_index(handle, trim(name) // c_null_char, i0))
    index_out = -1
    if (status == MKPP_SUCCESS) index_out = int(i0) + 1
  end subroutine mkpp_amd_species_index_f

  ! The handle's last error message.
  subroutine mkpp_amd_last_error_f(handle, message)
    type(c_ptr), intent(in) :: handle
    character(len=*), intent(out) :: message
    character(kind=c_char) :: buf(512)
    integer(c_int) :: rc
    integer :: i
    buf = c_null_char
    rc = mkpp_amd_get_last_error(handle, buf, 512_c_size_t)
    message = ''
    do i = 1, min(len(message), 512)
      if (buf(i) == c_null_char) exit
      message(i:i) = buf(i)
    end do
  end subroutine mkpp_amd_last_error_f

  ! p advanced by `bytes` (a device sub-array, e.g. the cells of one batch segment).
  function mkpp_amd_offset(p, bytes) result(q)
    type(c_ptr), intent(in) :: p
    integer(c_intptr_t), intent(in) :: bytes
    type(c_ptr) :: q
    q = transfer(transfer(p, 0_c_intptr_t) + bytes, q)
  end function mkpp_amd_offset

end module mkpp_amd_mod
