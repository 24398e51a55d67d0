// Drop-in library libmkpp_<MKPP_MECH>.so: the reference's unprefixed C ABI
// (src/mkpp/templates/host_api/mkpp.h.j2:36-72) bound to one mechanism,
// forwarding to libmkpp_amd.so.  Build with -DMKPP_MECH=\"<name>\".
#include "../../include/mkpp_amd.h"

#ifndef MKPP_MECH
#error "define MKPP_MECH to the mechanism name"
#endif

extern "C" {
int mkpp_create_handle(mkpp_handle_t* h, size_t n) { return mkpp_amd_create_handle(MKPP_MECH, h, n); }
int mkpp_destroy_handle(mkpp_handle_t h) { return mkpp_amd_destroy_handle(h); }
int mkpp_initialize(mkpp_handle_t h) { return mkpp_amd_initialize(h); }
int mkpp_get_species_count(mkpp_handle_t h, size_t* c) { return mkpp_amd_get_species_count(h, c); }
int mkpp_get_reaction_count(mkpp_handle_t h, size_t* c) { return mkpp_amd_get_reaction_count(h, c); }
int mkpp_get_photolysis_count(mkpp_handle_t h, size_t* c) { return mkpp_amd_get_photolysis_count(h, c); }
int mkpp_get_species_index(mkpp_handle_t h, const char* s, int* i) { return mkpp_amd_get_species_index(h, s, i); }
int mkpp_get_photolysis_index(mkpp_handle_t h, const char* s, int* i) { return mkpp_amd_get_photolysis_index(h, s, i); }
int mkpp_set_state_ptrs(mkpp_handle_t h, double* c, const double* t, const double* p, const double* r) {
  return mkpp_amd_set_state_ptrs(h, c, t, p, r);
}
int mkpp_set_rate_parameter_ptrs(mkpp_handle_t h, const double* v, size_t n) {
  return mkpp_amd_set_rate_parameter_ptrs(h, v, n);
}
int mkpp_set_photolysis_ptrs(mkpp_handle_t h, const double* v) { return mkpp_amd_set_photolysis_ptrs(h, v); }
int mkpp_set_cloud_water_ptr(mkpp_handle_t h, const double* v) { return mkpp_amd_set_cloud_water_ptr(h, v); }
int mkpp_set_aerosol_ptrs(mkpp_handle_t h, const double* rh, const double* sa) { return mkpp_amd_set_aerosol_ptrs(h, rh, sa); }
int mkpp_integrate(mkpp_handle_t h, double dt) { return mkpp_amd_integrate(h, dt); }
int mkpp_evaluate_rates(mkpp_handle_t h, double* out) { return mkpp_amd_evaluate_rates(h, out); }
int mkpp_evaluate_jacobian(mkpp_handle_t h, double* out) { return mkpp_amd_evaluate_jacobian(h, out); }
int mkpp_get_last_error(mkpp_handle_t h, char* b, size_t n) { return mkpp_amd_get_last_error(h, b, n); }
}
