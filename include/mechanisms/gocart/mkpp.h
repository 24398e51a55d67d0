/*
 * mkpp.h -- C ABI of the compiled mechanism gocart (libmkpp_gocart.so).
 * Generated by mkpp_amd.gen.emit_host; do not edit.
 *
 * The reference's generated C header (src/mkpp/templates/host_api/mkpp.h.j2:17-72):
 * the same symbols, argument meaning, LayoutLeft buffers and return codes,
 * and the mechanism's metadata macros.  The MKPP_DEFAULT_* macros state the
 * generation-time solver options this drop-in's handles start with.
 */
#ifndef MKPP_AMD_GENERATED_GOCART_H
#define MKPP_AMD_GENERATED_GOCART_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define MKPP_SUCCESS 0
#define MKPP_ERROR_NULL_POINTER -101
#define MKPP_ERROR_INVALID_HANDLE -102
#define MKPP_ERROR_INVALID_ARGUMENT -103
#define MKPP_ERROR_INVALID_STATE -104
#define MKPP_ERROR_CONVERGENCE -105
#define MKPP_ERROR_NOT_FOUND -106

typedef struct mkpp_handle_s* mkpp_handle_t;

/* Mechanism metadata (mkpp.h.j2:30-33). */
#define MKPP_NUM_RATE_PARAMETERS 1
#define MKPP_NUM_SPECIES 26
#define MKPP_NUM_REACTIONS 27
#define MKPP_NUM_PHOTOLYSIS 1

/* Generation-time solver options of this drop-in: method id (0 ROS-2,
 * 1 ROS-3, 2 ROS-4, 3 RODAS-3, 4 RODAS-4), tolerances and the first step as
 * a fraction of the chemistry interval. */
#define MKPP_DEFAULT_METHOD 1
#define MKPP_DEFAULT_ATOL 0.001
#define MKPP_DEFAULT_RTOL 1e-06
#define MKPP_DEFAULT_INITIAL_STEP_SEED 1e-06

int mkpp_create_handle(mkpp_handle_t* handle_out, size_t num_cells);
int mkpp_destroy_handle(mkpp_handle_t handle);
int mkpp_initialize(mkpp_handle_t handle);
int mkpp_get_species_count(mkpp_handle_t handle, size_t* count_out);
int mkpp_get_reaction_count(mkpp_handle_t handle, size_t* count_out);
int mkpp_get_photolysis_count(mkpp_handle_t handle, size_t* count_out);
int mkpp_get_species_index(mkpp_handle_t handle, const char* species_name, int* index_out);
int mkpp_get_photolysis_index(mkpp_handle_t handle, const char* photo_name, int* index_out);
int mkpp_set_state_ptrs(mkpp_handle_t handle, double* species_conc, const double* temp, const double* pressure,
                        const double* air_density);
/* LayoutLeft: values[cell + num_cells * parameter], MKPP_NUM_RATE_PARAMETERS parameters per cell. */
int mkpp_set_rate_parameter_ptrs(mkpp_handle_t handle, const double* values, size_t count);
int mkpp_set_photolysis_ptrs(mkpp_handle_t handle, const double* photolysis_rates);
int mkpp_set_cloud_water_ptr(mkpp_handle_t handle, const double* values);
int mkpp_set_aerosol_ptrs(mkpp_handle_t handle, const double* relative_humidity, const double* aerosol_surface_area);
int mkpp_integrate(mkpp_handle_t handle, double dt);
/* rates_out[cell + num_cells * species]; jac_out[cell + num_cells * (i + MKPP_NUM_SPECIES * j)] = dF_i/dC_j. */
int mkpp_evaluate_rates(mkpp_handle_t handle, double* rates_out);
int mkpp_evaluate_jacobian(mkpp_handle_t handle, double* jac_out);
int mkpp_get_last_error(mkpp_handle_t handle, char* buffer, size_t buffer_len);

#ifdef __cplusplus
}
#endif

#endif /* MKPP_AMD_GENERATED_GOCART_H */
