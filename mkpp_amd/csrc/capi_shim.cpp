// Drop-in library libmkpp_<MKPP_MECH>.so: the reference's unprefixed C ABI
// (src/mkpp/templates/host_api/mkpp.h.j2:36-72) bound to one mechanism,
// forwarding to libmkpp_amd.so.  Build with -DMKPP_MECH=\"<name>\".
//
// The reference bakes the solver (--solver) and the tolerances and first-step
// seed (the environment file's solver block) into each generated mechanism
// (src/mkpp/cli.py:229-236, 394-399; template_context.py:283-317).  A drop-in
// built with -DMKPP_DEFAULT_METHOD=<id>, -DMKPP_DEFAULT_ATOL=<a>,
// -DMKPP_DEFAULT_RTOL=<r> and -DMKPP_DEFAULT_INITIAL_STEP_SEED=<s> starts
// every handle with those options; without them it keeps the reference's
// shipped defaults (ROS-3, 1e-3, 1e-6, 1e-6), which are libmkpp_amd's.
#include <cstdlib>
#include <vector>

#include "../../include/mkpp_amd.h"

#ifndef MKPP_MECH
#error "define MKPP_MECH to the mechanism name"
#endif

namespace {
int apply_defaults(mkpp_handle_t h) {
#if defined(MKPP_DEFAULT_METHOD) || defined(MKPP_DEFAULT_ATOL) || defined(MKPP_DEFAULT_RTOL) || \
    defined(MKPP_DEFAULT_INITIAL_STEP_SEED)
#ifndef MKPP_DEFAULT_METHOD
#define MKPP_DEFAULT_METHOD MKPP_METHOD_ROS3
#endif
#ifndef MKPP_DEFAULT_ATOL
#define MKPP_DEFAULT_ATOL 1.0e-3
#endif
#ifndef MKPP_DEFAULT_RTOL
#define MKPP_DEFAULT_RTOL 1.0e-6
#endif
#ifndef MKPP_DEFAULT_INITIAL_STEP_SEED
#define MKPP_DEFAULT_INITIAL_STEP_SEED 1.0e-6
#endif
  size_t n = 0;
  int rc = mkpp_amd_get_species_count(h, &n);
  if (rc != MKPP_SUCCESS) return rc;
  std::vector<double> atol(n, MKPP_DEFAULT_ATOL), rtol(n, MKPP_DEFAULT_RTOL);
  return mkpp_amd_set_solver_options(h, MKPP_DEFAULT_METHOD, atol.data(), rtol.data(), MKPP_DEFAULT_INITIAL_STEP_SEED);
#else
  (void)h;
  return MKPP_SUCCESS;
#endif
}
}  // namespace

extern "C" {
int mkpp_create_handle(mkpp_handle_t* h, size_t n) {
  int rc = mkpp_amd_create_handle(MKPP_MECH, h, n);
  if (rc != MKPP_SUCCESS) return rc;
  if ((rc = apply_defaults(*h)) != MKPP_SUCCESS) {
    mkpp_amd_destroy_handle(*h);
    *h = nullptr;
  }
  return rc;
}
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
