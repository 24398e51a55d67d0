/*
 * mkpp_amd.h — C ABI of libmkpp_amd.so (every bundled mechanism in one library).
 *
 * Three groups of entry points:
 *
 * 1. Registry — the C face of the reference's MechanismRegistry
 *    (src/mkpp_host/include/mkpp_host/mechanism_manager.hpp:28-116:
 *    get_registered_mechanisms / has_mechanism / get_info).
 *
 * 2. Handle API — the reference's per-mechanism C ABI
 *    (src/mkpp/templates/host_api/mkpp.h.j2:36-72) with the mechanism chosen
 *    at handle creation: mkpp_amd_X(handle, ...) has exactly the meaning of
 *    mkpp_X(handle, ...) in mkpp.h.  Extensions without a reference
 *    counterpart are marked below.
 *
 * 3. Device API — stream-ordered entry points on device pointers (no staging,
 *    no allocation), for hosts that keep the chemistry state resident in HBM.
 *    No reference counterpart: the reference runs its C ABI on the host only
 *    (mkpp_c_api.cpp.j2:19) and defers an async API
 *    (PERFORMANCE_DESIGN_REVIEW.md:26, 68-70).
 *
 * Method ids: 0 ROS-2, 1 ROS-3, 2 ROS-4, 3 RODAS-3, 4 RODAS-4
 * (the reference's --solver choices, src/mkpp/cli.py:394-399).
 */
#ifndef MKPP_AMD_H
#define MKPP_AMD_H

#include <stddef.h>

#include "mkpp.h"

#ifdef __cplusplus
extern "C" {
#endif

#define MKPP_METHOD_ROS2 0
#define MKPP_METHOD_ROS3 1
#define MKPP_METHOD_ROS4 2
#define MKPP_METHOD_RODAS3 3
#define MKPP_METHOD_RODAS4 4

/* Per-cell status codes (src/mkpp_host/include/mkpp_host/dispatcher.hpp:15-21). */
#define MKPP_CELL_SUCCESS 0
#define MKPP_CELL_ERR_SINGULAR_MATRIX 1
#define MKPP_CELL_ERR_NON_FINITE_STATE 2
#define MKPP_CELL_ERR_CONVERGENCE_FAILURE 3
#define MKPP_CELL_ERR_NUMERICAL_FAILURE 4

/* ---- 1. registry (mechanism_manager.hpp:28-116) ---- */
int mkpp_amd_mechanism_count(void);
const char* mkpp_amd_mechanism_name(int index);
int mkpp_amd_has_mechanism(const char* name);
int mkpp_amd_mechanism_info(const char* name, int* num_species, int* num_reactions, int* num_photolysis,
                            int* num_rate_parameters, int* supports_runtime_environment);
int mkpp_amd_species_name(const char* mechanism, int index, const char** name_out);

/* ---- 2. handle API (mkpp.h.j2:36-72; mkpp_c_api.cpp.j2:72-371) ---- */
int mkpp_amd_create_handle(const char* mechanism, mkpp_handle_t* handle_out, size_t num_cells);
int mkpp_amd_destroy_handle(mkpp_handle_t handle);
int mkpp_amd_initialize(mkpp_handle_t handle);
int mkpp_amd_get_species_count(mkpp_handle_t handle, size_t* count_out);
int mkpp_amd_get_reaction_count(mkpp_handle_t handle, size_t* count_out);
int mkpp_amd_get_photolysis_count(mkpp_handle_t handle, size_t* count_out);
int mkpp_amd_get_species_index(mkpp_handle_t handle, const char* species_name, int* index_out);
int mkpp_amd_get_photolysis_index(mkpp_handle_t handle, const char* photo_name, int* index_out);
int mkpp_amd_set_state_ptrs(mkpp_handle_t handle, double* species_conc, const double* temp, const double* pressure,
                            const double* air_density);
int mkpp_amd_set_rate_parameter_ptrs(mkpp_handle_t handle, const double* values, size_t count);
int mkpp_amd_set_photolysis_ptrs(mkpp_handle_t handle, const double* photolysis_rates);
int mkpp_amd_set_cloud_water_ptr(mkpp_handle_t handle, const double* values);
int mkpp_amd_set_aerosol_ptrs(mkpp_handle_t handle, const double* relative_humidity,
                              const double* aerosol_surface_area);
int mkpp_amd_integrate(mkpp_handle_t handle, double dt);
int mkpp_amd_evaluate_rates(mkpp_handle_t handle, double* rates_out);
int mkpp_amd_evaluate_jacobian(mkpp_handle_t handle, double* jac_out);
int mkpp_amd_get_last_error(mkpp_handle_t handle, char* buffer, size_t buffer_len);

/* Extension (no reference counterpart; the reference bakes these into the
 * generated header: src/mkpp/template_context.py:283-340).  atol/rtol: NULL
 * keeps the current values, else num_species entries in physical order. */
int mkpp_amd_set_solver_options(mkpp_handle_t handle, int method, const double* atol, const double* rtol,
                                double initial_step_seed);
/* Extension: the handle's solver options (method id, num_species atol and
 * rtol values in physical order, seed); any pointer may be NULL. */
int mkpp_amd_get_solver_options(mkpp_handle_t handle, int* method, double* atol, double* rtol,
                                double* initial_step_seed);
/* Extension: attempted-step limit per integrate() call (reference: unbounded). */
int mkpp_amd_set_max_attempts(mkpp_handle_t handle, long max_attempts);
/* Extension: `steps` consecutive integrate(dt) calls per cell in one launch
 * (HostExecutionParams::steps, dispatcher.hpp:80-98). */
int mkpp_amd_integrate_steps(mkpp_handle_t handle, double dt, int steps);
/* Extension: per-cell status and accepted / attempted step counts of the last
 * integration (host arrays of num_cells; any may be NULL). */
int mkpp_amd_get_cell_status(mkpp_handle_t handle, int* status, int* accepted, int* attempted);

/* ---- 3. device API (no reference counterpart) ----
 * All pointers are device pointers; `stream` is a hipStream_t (NULL = default).
 * conc[cell + stride*species] (stride >= num_cells), rate_parameters
 * [cell + parameter_stride*slot] or NULL, temp/air_density [num_cells] or NULL
 * (300 K, 2.4476e19), atol/rtol [num_species], status/accepted/attempted
 * [num_cells] or NULL (accepted/attempted are incremented).  Returns after
 * enqueueing; allocates nothing. */
int mkpp_amd_device_integrate(const char* mechanism, double* conc, long num_cells, long stride, const double* temp,
                              const double* air_density, const double* rate_parameters, long parameter_stride,
                              double dt, int steps, int method, double initial_step_seed, const double* atol,
                              const double* rtol, int* status, int* accepted, int* attempted, long max_attempts,
                              void* stream);
/* The reference's integrate_with_reduction
 * (src/mkpp/templates/solver_variants/integrate_with_reduction.j2:51-306; the
 * reference compiles it under MKPP_ENABLE_REDUCTION and exposes no C entry
 * point for it).  Each attempted step freezes the species whose
 * |F| / (atol[t] + rtol[t] |y|) is below importance_threshold (t: the
 * species' position in the mechanism's solver permutation, as the reference
 * indexes the tolerance arrays there; pass uniform tolerances to make the
 * distinction moot).  Frozen species keep their value for that step.  dt
 * starts at min(dt, 1 s) and follows the reference's reduced controller (no
 * h_min, no rejection rules); there is no initial-step seed.  Same layout,
 * status and attempt limit as mkpp_amd_device_integrate. */
int mkpp_amd_device_integrate_reduced(const char* mechanism, double* conc, long num_cells, long stride,
                                      const double* temp, const double* air_density, const double* rate_parameters,
                                      long parameter_stride, double dt, int steps, int method,
                                      double importance_threshold, const double* atol, const double* rtol,
                                      int* status, int* accepted, int* attempted, long max_attempts, void* stream);
/* Validation backend: the same integration over the dense unrolled LU and
 * solves the reference compiles under MKPP_USE_UNROLLED_REFERENCE
 * (src/mkpp/templates/solver_variants/integrate.j2:136-141, 255-259;
 * device_impl/factorize.hpp.j2:13-18, device_impl/solve.hpp.j2:12-21): one
 * cell per lane, N x N matrices in the solver permutation.  Lane-per-cell
 * mechanisms only (MKPP_ERROR_INVALID_ARGUMENT otherwise); slow by design.
 * Arguments and results as mkpp_amd_device_integrate, whose results it equals
 * bit for bit. */
int mkpp_amd_device_integrate_unrolled(const char* mechanism, double* conc, long num_cells, long stride,
                                       const double* temp, const double* air_density, const double* rate_parameters,
                                       long parameter_stride, double dt, int steps, int method,
                                       double initial_step_seed, const double* atol, const double* rtol, int* status,
                                       int* accepted, int* attempted, long max_attempts, void* stream);
/* what: 0 = rates (out[cell + out_stride*species]),
 *       1 = dense Jacobian (out[cell + out_stride*(i + num_species*j)]). */
int mkpp_amd_device_evaluate(const char* mechanism, int what, const double* conc, long num_cells, long stride,
                             const double* temp, const double* air_density, const double* rate_parameters,
                             long parameter_stride, double* out, long out_stride, void* stream);

/* ---- 4. persistent batch over several mechanisms (replaces one launch per
 * mechanism: src/mkpp_host/include/mkpp_host/dispatcher.hpp:258-362 and
 * MechanismRegistry::execute, mechanism_manager.hpp:118-209) ----
 * One launch integrates every segment's cells; wavefronts take work items
 * (one cell of a wave-per-cell mechanism, 64*cells_per_lane cells of a
 * lane-per-cell one) from eight device queues, one per XCD, with stealing.
 * Segments may appear in any order; pointers and layouts are those of
 * mkpp_amd_device_integrate.  `workspace` is device memory of at least
 * MKPP_AMD_BATCH_WORKSPACE_BYTES, private to this call until it completes
 * (the queue heads; the call resets them on the stream, so it may be captured
 * in a graph).  workgroups = 0 launches as many as the device holds at once.
 * Allocates nothing. */
#define MKPP_AMD_BATCH_MAX_SEGMENTS 8
#define MKPP_AMD_BATCH_WORKSPACE_BYTES 32

typedef struct mkpp_amd_batch_segment {
  const char* mechanism;
  double* conc;
  long num_cells;
  long stride;
  const double* temp;
  const double* air_density;
  const double* rate_parameters;
  long parameter_stride;
  const double* atol;
  const double* rtol;
  int* status;
  int* accepted;
  int* attempted;
} mkpp_amd_batch_segment;

int mkpp_amd_device_integrate_batch(const mkpp_amd_batch_segment* segments, int num_segments, double dt, int steps,
                                    int method, double initial_step_seed, long max_attempts, int cells_per_lane,
                                    int workgroups, void* workspace, void* stream);
/* The same persistent batch with no workspace argument and no allocation.
 * No reference counterpart: for callers that cannot allocate device memory
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  The queue heads live in a static device array of
 * MKPP_AMD_BATCH_STATIC_SCRATCH_BYTES that each GPU the library runs on holds
 * until the process ends; the heads need MKPP_AMD_BATCH_WORKSPACE_BYTES of it
 * whatever the input size.  Every call on one GPU shares that GPU's array and
 * resets it on the stream before its launch, so calls that overlap on one GPU
 * (concurrent streams or host threads) race: serialize them.  Arguments,
 * limits and results are those of mkpp_amd_device_integrate_batch. */
#define MKPP_AMD_BATCH_STATIC_SCRATCH_BYTES 256
int mkpp_amd_device_integrate_batch_static(const mkpp_amd_batch_segment* segments, int num_segments, double dt,
                                           int steps, int method, double initial_step_seed, long max_attempts,
                                           int cells_per_lane, int workgroups, void* stream);
/* Workgroups of the batch kernel for `method` that one CU holds at once, and
 * the CU count of the current device (the default grid is their product). */
int mkpp_amd_batch_capacity(int method, int* per_cu, int* num_cus);

/* ---- 6. registry execution (MechanismRegistry::execute) ----
 * The reference's name-dispatched execute(name, state, HostExecutionParams)
 * (src/mkpp_host/include/mkpp_host/mechanism_manager.hpp:118-209) with the
 * per-cell semantics of its ParamsTiledCellIntegrator
 * (src/mkpp_host/include/mkpp_host/dispatcher.hpp:80-98, 197-255): per cell,
 * T = cell_temperature ? cell_temperature[cell] : temp, the air density
 * likewise, and the forcing is the cell's row of cell_jvals when it is given
 * (both extents non-zero), else the shared vector jvals, else zeros; the cell
 * is integrated `steps` times over dt.  The handle supplies the mechanism and
 * the solver options (mkpp_amd_set_solver_options; the per-mechanism drop-ins
 * preset their generation-time ones).  All pointers are device pointers;
 * conc is LayoutLeft (conc[cell + stride * species]).  The call is
 * synchronous, as the reference's execute fences: per-cell status and step
 * counts are then available from mkpp_amd_get_cell_status.  Returns
 * MKPP_ERROR_INVALID_ARGUMENT for an extent that is neither 0 nor num_cells,
 * a forcing shorter than the mechanism's runtime parameters, or a scalar
 * temperature or density that is not positive and finite where it is used;
 * MKPP_ERROR_CONVERGENCE if a cell reached the attempt limit.
 * clw and cell_cloud_water (cloud liquid water) gate the cloud-gated
 * heterogeneous reactions (gocart's SO2 + H2O2); rh and cell_humidity are
 * accepted and validated, and no bundled mechanism reads them (the
 * reference's generated code leaves RH unused as well). */
typedef struct mkpp_amd_execution_params {
  double dt;
  int steps;
  const double* jvals; /* [jvals_extent] shared by every cell, or NULL */
  long jvals_extent;
  double temp;         /* K, where cell_temperature is NULL */
  double air_density;  /* molec cm-3, where cell_density is NULL */
  double rh;
  double clw;
  const double* cell_temperature; /* [num_cells] or NULL */
  const double* cell_humidity;    /* [num_cells] or NULL */
  const double* cell_density;     /* [num_cells] or NULL */
  const double* cell_cloud_water; /* [num_cells] or NULL */
  /* per-cell forcing: cell_jvals[cell * cell_jvals_stride0 + slot * cell_jvals_stride1] */
  const double* cell_jvals;
  long cell_jvals_extent0;
  long cell_jvals_extent1;
  long cell_jvals_stride0;
  long cell_jvals_stride1;
} mkpp_amd_execution_params;

int mkpp_amd_execute(mkpp_handle_t handle, double* conc, long num_cells, long stride,
                     const mkpp_amd_execution_params* params, void* stream);

/* ---- 5. differentiation (every mechanism and method) ----
 * The lane-per-cell mechanisms run these entry points on the wave engine's
 * tables of the same mechanism: their checkpointed forward equals
 * mkpp_amd_device_integrate within rounding, not bit for bit.
 * Replaces the reference's frozen-Jacobian checkpoint / adjoint templates
 * (solver_variants/integrate_fwd_checkpoint.j2:37-295, integrate_adj.j2:50-156)
 * with the exact adjoint of the discrete Rosenbrock map over the accepted step
 * sequence of a forward run.  Rate constants are per flux group (reactions
 * with the same rate law and reactants share one); k arrays are
 * [num_groups x k_stride] per cell, k_stride >= num_cells.  Checkpoints are
 * max_steps * num_variable doubles per cell (ckpt) and max_steps doubles
 * (hs); nsteps[cell] receives the accepted step count.  ckpt[(cell *
 * max_steps + step) * num_variable + row] is the state at the start of that
 * accepted step, rows in the sparse LU's order (the mechanism's species
 * permutation, lowering.py:894-996, with the fixed species left out).  Device pointers;
 * enqueued on `stream`; allocates nothing. */
#define MKPP_CELL_ERR_CHECKPOINT_OVERFLOW 5

int mkpp_amd_diff_info(const char* mechanism, int* num_groups, int* num_variable);
/* k[cell + k_stride*group] from the rate laws at the cells' T, air density and
 * rate parameters. */
int mkpp_amd_device_rate_constants(const char* mechanism, const double* conc, long num_cells, long stride,
                                   const double* temp, const double* air_density, const double* rate_parameters,
                                   long parameter_stride, double* k, long k_stride, void* stream);
/* replay = 0: adaptive forward integration over dt (one interval) that records
 * every accepted step; replay = 1: integrates the recorded step sizes again
 * without error control (the frozen-step map).  k may be NULL (rate laws). */
int mkpp_amd_device_integrate_checkpointed(const char* mechanism, double* conc, long num_cells, long stride,
                                           const double* temp, const double* air_density,
                                           const double* rate_parameters, long parameter_stride, const double* k,
                                           long k_stride, double dt, int method, double initial_step_seed,
                                           const double* atol, const double* rtol, int* status, int* accepted,
                                           int* attempted, long max_attempts, double* ckpt, double* hs, int* nsteps,
                                           int max_steps, int replay, void* stream);
/* lambda[cell + stride*species]: in dL/dy at the end of the interval, out
 * dL/dy at its start (fixed species included); kbar[cell + k_stride*group]
 * out dL/dk.  conc supplies the fixed species (any time: they do not change). */
int mkpp_amd_device_adjoint(const char* mechanism, const double* conc, long num_cells, long stride,
                            const double* temp, const double* air_density, const double* rate_parameters,
                            long parameter_stride, const double* k, long k_stride, int method, const double* ckpt,
                            const double* hs, const int* nsteps, int max_steps, double* lambda, double* kbar,
                            void* stream);

/* Tangent-linear model along the recorded steps (forward-mode derivative of
 * the same frozen-step map the adjoint differentiates; the reference's
 * frozen-Jacobian counterpart is integrate_tlm.j2:33-215).  conc: in y(0),
 * out y(T) (the replay of the recorded steps, as
 * mkpp_amd_device_integrate_checkpointed with replay = 1 computes it);
 * dy[cell + stride*species]: in the direction dy(0) (fixed species
 * included), out dy(T); dk[cell + k_stride*group] the direction of the rate
 * constants, or NULL for 0.  k as for the adjoint (NULL: rate laws).
 * status (may be NULL): MKPP_CELL_ERR_CHECKPOINT_OVERFLOW for a step count
 * outside [0, max_steps].  Device pointers; enqueued on `stream`; allocates
 * nothing. */
int mkpp_amd_device_tangent_linear(const char* mechanism, double* conc, long num_cells, long stride,
                                   const double* temp, const double* air_density, const double* rate_parameters,
                                   long parameter_stride, const double* k, long k_stride, int method,
                                   const double* hs, const int* nsteps, int max_steps, double* dy, const double* dk,
                                   int* status, void* stream);

#ifdef __cplusplus
}
#endif

#endif /* MKPP_AMD_H */
