/*
 * mkpp.h — drop-in C ABI of one compiled mechanism (libmkpp_<mechanism>.so).
 *
 * Same symbols, argument meaning, layouts and return codes as the reference's
 * generated header (src/mkpp/templates/host_api/mkpp.h.j2:17-72, rendered per
 * mechanism as mkpp.h).  Like the reference, the names are not namespaced: link
 * one mechanism library per program, or use mkpp_amd.h, which serves every
 * bundled mechanism through one library.
 *
 * Buffers are borrowed host pointers in LayoutLeft order:
 *   species_conc[cell + num_cells * species]
 *   rate parameters[cell + num_cells * parameter]
 * mkpp_integrate is synchronous: it stages the buffers to the GPU, runs one
 * fused Rosenbrock kernel and copies the state back before returning.
 */
#ifndef MKPP_AMD_MKPP_H
#define MKPP_AMD_MKPP_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Return codes (mkpp.h.j2:17-24). */
#define MKPP_SUCCESS 0
#define MKPP_ERROR_NULL_POINTER -101
#define MKPP_ERROR_INVALID_HANDLE -102
#define MKPP_ERROR_INVALID_ARGUMENT -103
#define MKPP_ERROR_INVALID_STATE -104
#define MKPP_ERROR_CONVERGENCE -105
#define MKPP_ERROR_NOT_FOUND -106

typedef struct mkpp_handle_s* mkpp_handle_t;

/* Lifecycle (mkpp.h.j2:36-38; mkpp_c_api.cpp.j2:72-116). */
int mkpp_create_handle(mkpp_handle_t* handle_out, size_t num_cells);
int mkpp_destroy_handle(mkpp_handle_t handle);
int mkpp_initialize(mkpp_handle_t handle);

/* Metadata and name lookups (mkpp.h.j2:41-45; mkpp_c_api.cpp.j2:118-170).
 * Photolysis names are "J_1" ... "J_n" as in the reference's fallback names. */
int mkpp_get_species_count(mkpp_handle_t handle, size_t* count_out);
int mkpp_get_reaction_count(mkpp_handle_t handle, size_t* count_out);
int mkpp_get_photolysis_count(mkpp_handle_t handle, size_t* count_out);
int mkpp_get_species_index(mkpp_handle_t handle, const char* species_name, int* index_out);
int mkpp_get_photolysis_index(mkpp_handle_t handle, const char* photo_name, int* index_out);

/* Zero-copy host binding (mkpp.h.j2:48-62; mkpp_c_api.cpp.j2:172-249).
 * Temperature, pressure and air density must be positive and finite per cell. */
int mkpp_set_state_ptrs(mkpp_handle_t handle, double* species_conc, const double* temp, const double* pressure,
                        const double* air_density);
int mkpp_set_rate_parameter_ptrs(mkpp_handle_t handle, const double* values, size_t count);
int mkpp_set_photolysis_ptrs(mkpp_handle_t handle, const double* photolysis_rates);
int mkpp_set_cloud_water_ptr(mkpp_handle_t handle, const double* values);
int mkpp_set_aerosol_ptrs(mkpp_handle_t handle, const double* relative_humidity, const double* aerosol_surface_area);

/* Computation (mkpp.h.j2:65-68; mkpp_c_api.cpp.j2:251-362).
 * mkpp_evaluate_rates:    rates_out[cell + num_cells * species]
 * mkpp_evaluate_jacobian: jac_out[cell + num_cells * (i + num_species * j)] = dF_i/dC_j
 *                         (structural zeros are written as 0.0). */
int mkpp_integrate(mkpp_handle_t handle, double dt);
int mkpp_evaluate_rates(mkpp_handle_t handle, double* rates_out);
int mkpp_evaluate_jacobian(mkpp_handle_t handle, double* jac_out);

/* Diagnostics (mkpp.h.j2:71; mkpp_c_api.cpp.j2:364-371). */
int mkpp_get_last_error(mkpp_handle_t handle, char* buffer, size_t buffer_len);

#ifdef __cplusplus
}
#endif

#endif /* MKPP_AMD_MKPP_H */
