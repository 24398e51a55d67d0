// C-ABI of libmkpp_amd.so: mechanism registry, handle API and device API.
//
// The handle API restates the reference's generated C-ABI
// (src/mkpp/templates/host_api/mkpp_c_api.cpp.j2:40-371): borrowed LayoutLeft
// host pointers, the same validation and the same return codes.  The
// per-mechanism drop-in libraries (libmkpp_<mech>.so, capi_shim.cpp) export
// the reference's unprefixed mkpp_* names on top of these mkpp_amd_* ones.
// Where the reference runs a host Kokkos parallel_for, this runs one HIP
// launch on the handle's stream with staging copies in and out; the
// device-pointer entry points skip the staging.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/mkpp_amd.h"
#include "batch.hpp"
#include "runtime.hpp"

namespace mkpp {

#include "gen/registry_list.inc"  // declares and lists mkpp_ops_<mech>()

const MechOps* find_mechanism(const char* name) {
  if (!name) return nullptr;
  for (int i = 0; i < mechanism_count(); ++i)
    if (std::strcmp(mechanism_at(i)->name, name) == 0) return mechanism_at(i);
  return nullptr;
}

int mechanism_count() { return (int)(sizeof(kRegistry) / sizeof(kRegistry[0])); }
const MechOps* mechanism_at(int i) { return (i >= 0 && i < mechanism_count()) ? kRegistry[i]() : nullptr; }

}  // namespace mkpp

using mkpp::DeviceArgs;
using mkpp::EvalArgs;
using mkpp::MechOps;

struct mkpp_handle_s {
  const MechOps* ops = nullptr;
  size_t num_cells = 0;
  bool initialized = false;
  double* conc = nullptr;
  const double* temp = nullptr;
  const double* pressure = nullptr;
  const double* air_density = nullptr;
  const double* params = nullptr;
  const double* cloud_water = nullptr;
  const double* rh = nullptr;
  const double* aerosol_sa = nullptr;
  // solver options (defaults = the reference's shipped headers: ROS-3,
  // atol 1e-3, rtol 1e-6, seed 1e-6; template_context.py:283-317)
  int method = mkpp::ROS3;
  double seed = 1.0e-6;
  std::vector<double> atol, rtol;
  long max_attempts = 100000;  // KPP Rosenbrock default Max_no_steps
  // device staging
  hipStream_t stream = nullptr;
  size_t cap_cells = 0;
  double *d_conc = nullptr, *d_temp = nullptr, *d_rho = nullptr, *d_par = nullptr, *d_tol = nullptr, *d_out = nullptr;
  double* d_clw = nullptr;  // staged cloud liquid water (when a cloud-water pointer is set)
  size_t cap_out = 0;
  int *d_status = nullptr, *d_acc = nullptr, *d_att = nullptr;
  std::vector<int> status, accepted, attempted;
  std::string last_error;

  void set_error(const std::string& s) { last_error = s; }
  ~mkpp_handle_s() {
    for (void* p : {(void*)d_conc, (void*)d_temp, (void*)d_rho, (void*)d_par, (void*)d_tol, (void*)d_out, (void*)d_clw,
                    (void*)d_status, (void*)d_acc, (void*)d_att})
      (void)hipFree(p);
    if (stream) (void)hipStreamDestroy(stream);
  }
};

namespace {

int hip_fail(mkpp_handle_s* h, hipError_t e, const char* what) {
  if (h) h->set_error(std::string("FATAL ERROR: ") + what + ": " + hipGetErrorString(e));
  return MKPP_ERROR_INVALID_STATE;
}

#define HIPCHK(h, call)                                   \
  do {                                                    \
    hipError_t e_ = (call);                               \
    if (e_ != hipSuccess) return hip_fail((h), e_, #call); \
  } while (0)

// mkpp_c_api.cpp.j2:40-66
int validate_state(mkpp_handle_s* h) {
  if (!h->conc) {
    h->set_error("FATAL ERROR: species concentration pointer is not bound");
    return MKPP_ERROR_NULL_POINTER;
  }
  if (h->num_cells == 0) {
    h->set_error("FATAL ERROR: species concentration view extents do not match the handle and mechanism");
    return MKPP_ERROR_INVALID_STATE;
  }
  if (!h->temp || !h->air_density || !h->pressure) {
    h->set_error("FATAL ERROR: per-cell meteorology is not bound");
    return MKPP_ERROR_NULL_POINTER;
  }
  for (size_t c = 0; c < h->num_cells; ++c) {
    if (!std::isfinite(h->temp[c]) || h->temp[c] <= 0.0 || !std::isfinite(h->air_density[c]) ||
        h->air_density[c] <= 0.0 || !std::isfinite(h->pressure[c]) || h->pressure[c] <= 0.0) {
      h->set_error("FATAL ERROR: cell temperature, density and pressure must be positive and finite");
      return MKPP_ERROR_INVALID_ARGUMENT;
    }
  }
  return MKPP_SUCCESS;
}

template <class T>
int grow(mkpp_handle_s* h, T** p, size_t n) {
  (void)hipFree(*p);
  *p = nullptr;
  HIPCHK(h, hipMalloc((void**)p, n * sizeof(T) + 64));
  return MKPP_SUCCESS;
}

// Allocate device staging for the handle's cell count (once).
int ensure_device(mkpp_handle_s* h) {
  if (!h->stream) HIPCHK(h, hipStreamCreateWithFlags(&h->stream, hipStreamNonBlocking));
  if (h->cap_cells >= h->num_cells) return MKPP_SUCCESS;
  const size_t n = h->num_cells, N = (size_t)h->ops->num_species;
  const size_t P = (size_t)(h->ops->num_runtime_parameters > 0 ? h->ops->num_runtime_parameters : 1);
  int rc;
  if ((rc = grow(h, &h->d_conc, n * N))) return rc;
  if ((rc = grow(h, &h->d_temp, n))) return rc;
  if ((rc = grow(h, &h->d_rho, n))) return rc;
  if ((rc = grow(h, &h->d_clw, n))) return rc;
  if ((rc = grow(h, &h->d_par, n * P))) return rc;
  if ((rc = grow(h, &h->d_tol, 2 * N))) return rc;
  if ((rc = grow(h, &h->d_status, n))) return rc;
  if ((rc = grow(h, &h->d_acc, n))) return rc;
  if ((rc = grow(h, &h->d_att, n))) return rc;
  h->cap_cells = n;
  return MKPP_SUCCESS;
}

int upload_state(mkpp_handle_s* h) {
  const size_t n = h->num_cells, N = (size_t)h->ops->num_species;
  const size_t P = (size_t)h->ops->num_runtime_parameters;
  HIPCHK(h, hipMemcpyAsync(h->d_conc, h->conc, n * N * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->d_temp, h->temp, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->d_rho, h->air_density, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  if (P > 0 && h->params)
    HIPCHK(h, hipMemcpyAsync(h->d_par, h->params, n * P * sizeof(double), hipMemcpyHostToDevice, h->stream));
  if (h->cloud_water)
    HIPCHK(h, hipMemcpyAsync(h->d_clw, h->cloud_water, n * sizeof(double), hipMemcpyHostToDevice, h->stream));
  return MKPP_SUCCESS;
}

mkpp_handle_s* H(mkpp_handle_t h) { return reinterpret_cast<mkpp_handle_s*>(h); }

}  // namespace

extern "C" {

/* ---------------- registry ---------------- */
int mkpp_amd_mechanism_count(void) { return mkpp::mechanism_count(); }

const char* mkpp_amd_mechanism_name(int index) {
  const MechOps* o = mkpp::mechanism_at(index);
  return o ? o->name : nullptr;
}

int mkpp_amd_has_mechanism(const char* name) { return mkpp::find_mechanism(name) != nullptr; }

int mkpp_amd_mechanism_info(const char* name, int* num_species, int* num_reactions, int* num_photolysis,
                            int* num_rate_parameters, int* supports_runtime_environment) {
  const MechOps* o = mkpp::find_mechanism(name);
  if (!name) return MKPP_ERROR_NULL_POINTER;
  if (!o) return MKPP_ERROR_NOT_FOUND;
  if (num_species) *num_species = o->num_species;
  if (num_reactions) *num_reactions = o->num_reactions;
  if (num_photolysis) *num_photolysis = o->num_photolysis;
  if (num_rate_parameters) *num_rate_parameters = o->num_runtime_parameters;
  if (supports_runtime_environment) *supports_runtime_environment = o->supports_runtime_environment ? 1 : 0;
  return MKPP_SUCCESS;
}

int mkpp_amd_species_name(const char* mechanism, int index, const char** name_out) {
  const MechOps* o = mkpp::find_mechanism(mechanism);
  if (!mechanism || !name_out) return MKPP_ERROR_NULL_POINTER;
  if (!o) return MKPP_ERROR_NOT_FOUND;
  if (index < 0 || index >= o->num_species) return MKPP_ERROR_INVALID_ARGUMENT;
  *name_out = o->species_names[index];
  return MKPP_SUCCESS;
}

/* ---------------- handle lifecycle ---------------- */
int mkpp_amd_create_handle(const char* mechanism, mkpp_handle_t* handle_out, size_t num_cells) {
  if (!handle_out || !mechanism) return MKPP_ERROR_NULL_POINTER;
  if (num_cells == 0) return MKPP_ERROR_INVALID_ARGUMENT;
  const MechOps* o = mkpp::find_mechanism(mechanism);
  if (!o) return MKPP_ERROR_NOT_FOUND;
  try {
    auto* h = new mkpp_handle_s();
    h->ops = o;
    h->num_cells = num_cells;
    h->atol.assign(o->num_species, 1.0e-3);
    h->rtol.assign(o->num_species, 1.0e-6);
    *handle_out = reinterpret_cast<mkpp_handle_t>(h);
    return MKPP_SUCCESS;
  } catch (...) {
    return MKPP_ERROR_INVALID_STATE;
  }
}

int mkpp_amd_destroy_handle(mkpp_handle_t handle) {
  if (!handle) return MKPP_ERROR_NULL_POINTER;
  delete H(handle);
  return MKPP_SUCCESS;
}

int mkpp_amd_initialize(mkpp_handle_t handle) {
  if (!handle) return MKPP_ERROR_NULL_POINTER;
  H(handle)->initialized = true;
  return MKPP_SUCCESS;
}

int mkpp_amd_get_species_count(mkpp_handle_t handle, size_t* count_out) {
  if (!handle || !count_out) return MKPP_ERROR_NULL_POINTER;
  *count_out = (size_t)H(handle)->ops->num_species;
  return MKPP_SUCCESS;
}

int mkpp_amd_get_reaction_count(mkpp_handle_t handle, size_t* count_out) {
  if (!handle || !count_out) return MKPP_ERROR_NULL_POINTER;
  *count_out = (size_t)H(handle)->ops->num_reactions;
  return MKPP_SUCCESS;
}

int mkpp_amd_get_photolysis_count(mkpp_handle_t handle, size_t* count_out) {
  if (!handle || !count_out) return MKPP_ERROR_NULL_POINTER;
  *count_out = (size_t)H(handle)->ops->num_photolysis;
  return MKPP_SUCCESS;
}

int mkpp_amd_get_species_index(mkpp_handle_t handle, const char* species_name, int* index_out) {
  if (!handle || !species_name || !index_out) return MKPP_ERROR_NULL_POINTER;
  const MechOps* o = H(handle)->ops;
  for (int i = 0; i < o->num_species; ++i)
    if (std::strcmp(o->species_names[i], species_name) == 0) {
      *index_out = i;
      return MKPP_SUCCESS;
    }
  *index_out = -1;
  return MKPP_ERROR_NOT_FOUND;
}

// The reference's photolysis name map falls back to "J_<1-based index>"
// because its photolysis records carry no name (mkpp_c_api.cpp.j2:31-37).
int mkpp_amd_get_photolysis_index(mkpp_handle_t handle, const char* photo_name, int* index_out) {
  if (!handle || !photo_name || !index_out) return MKPP_ERROR_NULL_POINTER;
  const MechOps* o = H(handle)->ops;
  char buf[32];
  for (int i = 0; i < o->num_photolysis; ++i) {
    std::snprintf(buf, sizeof buf, "J_%d", i + 1);
    if (std::strcmp(buf, photo_name) == 0) {
      *index_out = i;
      return MKPP_SUCCESS;
    }
  }
  *index_out = -1;
  return MKPP_ERROR_NOT_FOUND;
}

/* ---------------- host memory binding ---------------- */
int mkpp_amd_set_state_ptrs(mkpp_handle_t handle, double* species_conc, const double* temp, const double* pressure,
                            const double* air_density) {
  if (!handle) return MKPP_ERROR_NULL_POINTER;
  mkpp_handle_s* h = H(handle);
  if (!species_conc || !temp || !pressure || !air_density) {
    h->set_error("FATAL ERROR: state binding requires concentration and meteorology pointers");
    return MKPP_ERROR_NULL_POINTER;
  }
  h->conc = species_conc;
  h->temp = temp;
  h->pressure = pressure;
  h->air_density = air_density;
  return validate_state(h);
}

int mkpp_amd_set_rate_parameter_ptrs(mkpp_handle_t handle, const double* values, size_t count) {
  if (!handle || !values) return MKPP_ERROR_NULL_POINTER;
  mkpp_handle_s* h = H(handle);
  if (count != (size_t)h->ops->num_runtime_parameters) {
    h->set_error("FATAL ERROR: runtime parameter count does not match the generated mechanism");
    return MKPP_ERROR_INVALID_ARGUMENT;
  }
  h->params = values;
  return MKPP_SUCCESS;
}

int mkpp_amd_set_photolysis_ptrs(mkpp_handle_t handle, const double* photolysis_rates) {
  if (!handle || !photolysis_rates) return MKPP_ERROR_NULL_POINTER;
  mkpp_handle_s* h = H(handle);
  if (h->ops->num_runtime_parameters > h->ops->num_photolysis) {
    h->set_error("FATAL ERROR: use mkpp_set_rate_parameter_ptrs for the complete runtime parameter array");
    return MKPP_ERROR_INVALID_ARGUMENT;
  }
  return mkpp_amd_set_rate_parameter_ptrs(handle, photolysis_rates, (size_t)h->ops->num_photolysis);
}

int mkpp_amd_set_cloud_water_ptr(mkpp_handle_t handle, const double* values) {
  if (!handle || !values) return MKPP_ERROR_NULL_POINTER;
  H(handle)->cloud_water = values;
  return MKPP_SUCCESS;
}

int mkpp_amd_set_aerosol_ptrs(mkpp_handle_t handle, const double* relative_humidity, const double* aerosol_surface_area) {
  if (!handle) return MKPP_ERROR_NULL_POINTER;
  if (relative_humidity) H(handle)->rh = relative_humidity;
  if (aerosol_surface_area) H(handle)->aerosol_sa = aerosol_surface_area;
  return MKPP_SUCCESS;
}

/* ---------------- solver options (no reference counterpart) ---------------- */
int mkpp_amd_set_solver_options(mkpp_handle_t handle, int method, const double* atol, const double* rtol,
                                double initial_step_seed) {
  if (!handle) return MKPP_ERROR_NULL_POINTER;
  mkpp_handle_s* h = H(handle);
  if (method < mkpp::ROS2 || method > mkpp::RODAS4) {
    h->set_error("FATAL ERROR: unknown Rosenbrock method");
    return MKPP_ERROR_INVALID_ARGUMENT;
  }
  if (!std::isfinite(initial_step_seed) || initial_step_seed <= 0.0) {
    h->set_error("FATAL ERROR: initial_step_seed must be a finite positive fraction of the chemistry interval");
    return MKPP_ERROR_INVALID_ARGUMENT;
  }
  h->method = method;
  h->seed = initial_step_seed;
  if (atol) h->atol.assign(atol, atol + h->ops->num_species);
  if (rtol) h->rtol.assign(rtol, rtol + h->ops->num_species);
  return MKPP_SUCCESS;
}

int mkpp_amd_get_solver_options(mkpp_handle_t handle, int* method, double* atol, double* rtol,
                                double* initial_step_seed) {
  if (!handle) return MKPP_ERROR_NULL_POINTER;
  const mkpp_handle_s* h = H(handle);
  if (method) *method = h->method;
  if (atol) std::copy(h->atol.begin(), h->atol.end(), atol);
  if (rtol) std::copy(h->rtol.begin(), h->rtol.end(), rtol);
  if (initial_step_seed) *initial_step_seed = h->seed;
  return MKPP_SUCCESS;
}

int mkpp_amd_set_max_attempts(mkpp_handle_t handle, long max_attempts) {
  if (!handle) return MKPP_ERROR_NULL_POINTER;
  H(handle)->max_attempts = max_attempts;
  return MKPP_SUCCESS;
}

/* ---------------- computation ---------------- */
int mkpp_amd_integrate_steps(mkpp_handle_t handle, double dt, int steps) {
  if (!handle) return MKPP_ERROR_NULL_POINTER;
  mkpp_handle_s* h = H(handle);
  if (!h->initialized) {
    h->set_error("FATAL ERROR: MKPP handle must be initialized before integration");
    return MKPP_ERROR_INVALID_STATE;
  }
  int rc = validate_state(h);
  if (rc != MKPP_SUCCESS) return rc;
  if (!std::isfinite(dt) || dt <= 0.0 || steps < 1) return MKPP_ERROR_INVALID_ARGUMENT;
  if ((rc = ensure_device(h)) != MKPP_SUCCESS) return rc;
  if ((rc = upload_state(h)) != MKPP_SUCCESS) return rc;
  const size_t n = h->num_cells, N = (size_t)h->ops->num_species;
  std::vector<double> tol(2 * N);
  for (size_t i = 0; i < N; ++i) {
    tol[i] = h->atol[i];
    tol[N + i] = h->rtol[i];
  }
  HIPCHK(h, hipMemcpyAsync(h->d_tol, tol.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice, h->stream));
  HIPCHK(h, hipMemsetAsync(h->d_acc, 0, n * sizeof(int), h->stream));
  HIPCHK(h, hipMemsetAsync(h->d_att, 0, n * sizeof(int), h->stream));
  DeviceArgs a{};
  a.conc = h->d_conc;
  a.ncells = (long)n;
  a.stride = (long)n;
  a.temp = h->d_temp;
  a.air_density = h->d_rho;
  a.jvals = (h->params && h->ops->num_runtime_parameters > 0) ? h->d_par : nullptr;
  a.jstride = (long)n;
  a.dt = dt;
  a.steps = steps;
  a.method = h->method;
  a.seed = h->seed;
  a.atol = h->d_tol;
  a.rtol = h->d_tol + N;
  a.status = h->d_status;
  a.accepted = h->d_acc;
  a.attempted = h->d_att;
  a.max_attempts = h->max_attempts;
  a.cloud_water = h->cloud_water ? h->d_clw : nullptr;
  if (h->ops->integrate(a, h->stream) != 0) {
    h->set_error("FATAL ERROR: integrate kernel launch failed");
    return MKPP_ERROR_INVALID_STATE;
  }
  h->status.resize(n);
  h->accepted.resize(n);
  h->attempted.resize(n);
  HIPCHK(h, hipMemcpyAsync(h->conc, h->d_conc, n * N * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->status.data(), h->d_status, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->accepted.data(), h->d_acc, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipMemcpyAsync(h->attempted.data(), h->d_att, n * sizeof(int), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  for (size_t c = 0; c < n; ++c) {
    if (h->status[c] == mkpp::ERR_CONVERGENCE_FAILURE) {
      h->set_error("FATAL ERROR: step-size control exceeded the attempted-step limit in cell " + std::to_string(c));
      return MKPP_ERROR_CONVERGENCE;
    }
  }
  return MKPP_SUCCESS;
}

int mkpp_amd_integrate(mkpp_handle_t handle, double dt) { return mkpp_amd_integrate_steps(handle, dt, 1); }

// MechanismRegistry::execute semantics on device pointers
// (mechanism_manager.hpp:118-209, dispatcher.hpp:197-255).
int mkpp_amd_execute(mkpp_handle_t handle, double* conc, long num_cells, long stride,
                     const mkpp_amd_execution_params* params, void* stream) {
  if (!handle) return MKPP_ERROR_NULL_POINTER;
  mkpp_handle_s* h = H(handle);
  if (!conc || !params) {
    h->set_error("FATAL ERROR: execute requires a state pointer and execution parameters");
    return MKPP_ERROR_NULL_POINTER;
  }
  if (!h->initialized) {
    h->set_error("FATAL ERROR: MKPP handle must be initialized before integration");
    return MKPP_ERROR_INVALID_STATE;
  }
  const mkpp_amd_execution_params& p = *params;
  const long P = h->ops->num_runtime_parameters;
  auto bad = [&](const char* msg) {
    h->set_error(std::string("FATAL ERROR: ") + msg);
    return MKPP_ERROR_INVALID_ARGUMENT;
  };
  if (num_cells != (long)h->num_cells || stride < num_cells)
    return bad("state view extents do not match the handle and mechanism");
  if (!std::isfinite(p.dt) || p.dt <= 0.0 || p.steps < 1) return bad("dt must be positive and finite, steps >= 1");
  const bool per_cell_forcing = p.cell_jvals && p.cell_jvals_extent0 != 0 && p.cell_jvals_extent1 != 0;
  if (p.cell_jvals_extent0 != 0 && p.cell_jvals_extent0 != num_cells)
    return bad("per-cell forcing extent does not match state");
  if (per_cell_forcing && p.cell_jvals_extent1 < P) return bad("per-cell forcing has too few runtime parameters");
  if (per_cell_forcing && (p.cell_jvals_stride0 < 0 || p.cell_jvals_stride1 < 0))
    return bad("per-cell forcing strides must not be negative");
  if (!per_cell_forcing && p.jvals && P > 0 && p.jvals_extent < P)
    return bad("photolysis array has fewer entries than the mechanism's runtime parameters");
  if (!p.cell_temperature && (!std::isfinite(p.temp) || p.temp <= 0.0))
    return bad("temperature must be positive and finite");
  if (!p.cell_density && (!std::isfinite(p.air_density) || p.air_density <= 0.0))
    return bad("air density must be positive and finite");
  if (!std::isfinite(p.rh) || !std::isfinite(p.clw)) return bad("relative humidity and cloud water must be finite");
  int rc;
  if ((rc = ensure_device(h)) != MKPP_SUCCESS) return rc;
  const size_t n = h->num_cells, N = (size_t)h->ops->num_species;
  std::vector<double> tol(2 * N);
  for (size_t i = 0; i < N; ++i) {
    tol[i] = h->atol[i];
    tol[N + i] = h->rtol[i];
  }
  hipStream_t s = stream ? (hipStream_t)stream : h->stream;
  HIPCHK(h, hipMemcpyAsync(h->d_tol, tol.data(), 2 * N * sizeof(double), hipMemcpyHostToDevice, s));
  HIPCHK(h, hipMemsetAsync(h->d_acc, 0, n * sizeof(int), s));
  HIPCHK(h, hipMemsetAsync(h->d_att, 0, n * sizeof(int), s));
  DeviceArgs a{};
  a.conc = conc;
  a.ncells = num_cells;
  a.stride = stride;
  a.temp = p.cell_temperature;
  a.temp_default = p.temp;
  a.air_density = p.cell_density;
  a.rho_default = p.air_density;
  if (P > 0 && per_cell_forcing) {
    a.jvals = p.cell_jvals;
    a.jcstride = p.cell_jvals_stride0;
    a.jstride = p.cell_jvals_stride1;
  } else if (P > 0 && p.jvals) {
    a.jvals = p.jvals;
    a.jcstride = 0;
    a.jstride = 1;
  }
  a.dt = p.dt;
  a.steps = p.steps;
  a.method = h->method;
  a.seed = h->seed;
  a.atol = h->d_tol;
  a.rtol = h->d_tol + N;
  a.status = h->d_status;
  a.accepted = h->d_acc;
  a.attempted = h->d_att;
  a.max_attempts = h->max_attempts;
  a.cloud_water = p.cell_cloud_water;
  a.clw_default = p.clw;
  if (h->ops->integrate(a, s) != 0) {
    h->set_error("FATAL ERROR: integrate kernel launch failed");
    return MKPP_ERROR_INVALID_STATE;
  }
  h->status.resize(n);
  h->accepted.resize(n);
  h->attempted.resize(n);
  HIPCHK(h, hipMemcpyAsync(h->status.data(), h->d_status, n * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipMemcpyAsync(h->accepted.data(), h->d_acc, n * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipMemcpyAsync(h->attempted.data(), h->d_att, n * sizeof(int), hipMemcpyDeviceToHost, s));
  HIPCHK(h, hipStreamSynchronize(s));
  for (size_t c = 0; c < n; ++c) {
    if (h->status[c] == mkpp::ERR_CONVERGENCE_FAILURE) {
      h->set_error("FATAL ERROR: step-size control exceeded the attempted-step limit in cell " + std::to_string(c));
      return MKPP_ERROR_CONVERGENCE;
    }
  }
  return MKPP_SUCCESS;
}

int mkpp_amd_get_cell_status(mkpp_handle_t handle, int* status, int* accepted, int* attempted) {
  if (!handle) return MKPP_ERROR_NULL_POINTER;
  mkpp_handle_s* h = H(handle);
  if (h->status.size() != h->num_cells) {
    h->set_error("FATAL ERROR: no integration has run on this handle");
    return MKPP_ERROR_INVALID_STATE;
  }
  if (status) std::memcpy(status, h->status.data(), h->num_cells * sizeof(int));
  if (accepted) std::memcpy(accepted, h->accepted.data(), h->num_cells * sizeof(int));
  if (attempted) std::memcpy(attempted, h->attempted.data(), h->num_cells * sizeof(int));
  return MKPP_SUCCESS;
}

static int evaluate(mkpp_handle_t handle, double* out, bool jac) {
  if (!handle) return MKPP_ERROR_NULL_POINTER;
  mkpp_handle_s* h = H(handle);
  if (!out) {
    h->set_error(jac ? "FATAL ERROR: Jacobian output pointer is null" : "FATAL ERROR: rates output pointer is null");
    return MKPP_ERROR_NULL_POINTER;
  }
  if (!h->initialized) {
    h->set_error(jac ? "FATAL ERROR: MKPP handle must be initialized before Jacobian evaluation"
                     : "FATAL ERROR: MKPP handle must be initialized before rates evaluation");
    return MKPP_ERROR_INVALID_STATE;
  }
  int rc = validate_state(h);
  if (rc != MKPP_SUCCESS) return rc;
  if ((rc = ensure_device(h)) != MKPP_SUCCESS) return rc;
  if ((rc = upload_state(h)) != MKPP_SUCCESS) return rc;
  const size_t n = h->num_cells, N = (size_t)h->ops->num_species;
  const size_t nout = jac ? n * N * N : n * N;
  if (h->cap_out < nout) {
    if ((rc = grow(h, &h->d_out, nout))) return rc;
    h->cap_out = nout;
  }
  EvalArgs e{};
  e.conc = h->d_conc;
  e.ncells = (long)n;
  e.stride = (long)n;
  e.temp = h->d_temp;
  e.air_density = h->d_rho;
  e.jvals = (h->params && h->ops->num_runtime_parameters > 0) ? h->d_par : nullptr;
  e.jstride = (long)n;
  e.out = h->d_out;
  e.out_stride = (long)n;
  e.cloud_water = h->cloud_water ? h->d_clw : nullptr;
  if ((jac ? h->ops->eval_jacobian(e, h->stream) : h->ops->eval_rates(e, h->stream)) != 0) {
    h->set_error("FATAL ERROR: evaluation kernel launch failed");
    return MKPP_ERROR_INVALID_STATE;
  }
  HIPCHK(h, hipMemcpyAsync(out, h->d_out, nout * sizeof(double), hipMemcpyDeviceToHost, h->stream));
  HIPCHK(h, hipStreamSynchronize(h->stream));
  return MKPP_SUCCESS;
}

int mkpp_amd_evaluate_rates(mkpp_handle_t handle, double* rates_out) { return evaluate(handle, rates_out, false); }
int mkpp_amd_evaluate_jacobian(mkpp_handle_t handle, double* jac_out) { return evaluate(handle, jac_out, true); }

int mkpp_amd_get_last_error(mkpp_handle_t handle, char* buffer, size_t buffer_len) {
  if (!handle || !buffer || buffer_len == 0) return MKPP_ERROR_NULL_POINTER;
  std::strncpy(buffer, H(handle)->last_error.c_str(), buffer_len - 1);
  buffer[buffer_len - 1] = '\0';
  return MKPP_SUCCESS;
}

/* ---------------- device-pointer API ---------------- */
int mkpp_amd_device_integrate(const char* mechanism, double* conc, long num_cells, long stride, const double* temp,
                              const double* air_density, const double* rate_parameters, long parameter_stride,
                              double dt, int steps, int method, double initial_step_seed, const double* atol,
                              const double* rtol, int* status, int* accepted, int* attempted, long max_attempts,
                              void* stream) {
  const MechOps* o = mkpp::find_mechanism(mechanism);
  if (!mechanism || !conc || !atol || !rtol) return MKPP_ERROR_NULL_POINTER;
  if (!o) return MKPP_ERROR_NOT_FOUND;
  if (num_cells < 0 || stride < num_cells || !std::isfinite(dt) || dt <= 0.0 || steps < 1 ||
      method < mkpp::ROS2 || method > mkpp::RODAS4 || !(initial_step_seed > 0.0))
    return MKPP_ERROR_INVALID_ARGUMENT;
  if (rate_parameters && parameter_stride < num_cells) return MKPP_ERROR_INVALID_ARGUMENT;
  DeviceArgs a{};
  a.conc = conc;
  a.ncells = num_cells;
  a.stride = stride;
  a.temp = temp;
  a.air_density = air_density;
  a.jvals = rate_parameters;
  a.jstride = parameter_stride;
  a.dt = dt;
  a.steps = steps;
  a.method = method;
  a.seed = initial_step_seed;
  a.atol = atol;
  a.rtol = rtol;
  a.status = status;
  a.accepted = accepted;
  a.attempted = attempted;
  a.max_attempts = max_attempts;
  return o->integrate(a, (hipStream_t)stream) == 0 ? MKPP_SUCCESS : MKPP_ERROR_INVALID_STATE;
}

int mkpp_amd_device_integrate_unrolled(const char* mechanism, double* conc, long num_cells, long stride,
                                       const double* temp, const double* air_density, const double* rate_parameters,
                                       long parameter_stride, double dt, int steps, int method,
                                       double initial_step_seed, const double* atol, const double* rtol, int* status,
                                       int* accepted, int* attempted, long max_attempts, void* stream) {
  const MechOps* o = mkpp::find_mechanism(mechanism);
  if (!mechanism || !conc || !atol || !rtol) return MKPP_ERROR_NULL_POINTER;
  if (!o) return MKPP_ERROR_NOT_FOUND;
  if (num_cells < 0 || stride < num_cells || !std::isfinite(dt) || dt <= 0.0 || steps < 1 ||
      method < mkpp::ROS2 || method > mkpp::RODAS4 || !(initial_step_seed > 0.0))
    return MKPP_ERROR_INVALID_ARGUMENT;
  if (rate_parameters && parameter_stride < num_cells) return MKPP_ERROR_INVALID_ARGUMENT;
  if (!o->integrate_unrolled) return MKPP_ERROR_INVALID_ARGUMENT;
  DeviceArgs a{};
  a.conc = conc;
  a.ncells = num_cells;
  a.stride = stride;
  a.temp = temp;
  a.air_density = air_density;
  a.jvals = rate_parameters;
  a.jstride = parameter_stride;
  a.dt = dt;
  a.steps = steps;
  a.method = method;
  a.seed = initial_step_seed;
  a.atol = atol;
  a.rtol = rtol;
  a.status = status;
  a.accepted = accepted;
  a.attempted = attempted;
  a.max_attempts = max_attempts;
  return o->integrate_unrolled(a, (hipStream_t)stream) == 0 ? MKPP_SUCCESS : MKPP_ERROR_INVALID_STATE;
}

int mkpp_amd_device_integrate_reduced(const char* mechanism, double* conc, long num_cells, long stride,
                                      const double* temp, const double* air_density, const double* rate_parameters,
                                      long parameter_stride, double dt, int steps, int method,
                                      double importance_threshold, const double* atol, const double* rtol,
                                      int* status, int* accepted, int* attempted, long max_attempts, void* stream) {
  const MechOps* o = mkpp::find_mechanism(mechanism);
  if (!mechanism || !conc || !atol || !rtol) return MKPP_ERROR_NULL_POINTER;
  if (!o) return MKPP_ERROR_NOT_FOUND;
  if (num_cells < 0 || stride < num_cells || !std::isfinite(dt) || dt <= 0.0 || steps < 1 ||
      method < mkpp::ROS2 || method > mkpp::RODAS4 || std::isnan(importance_threshold))
    return MKPP_ERROR_INVALID_ARGUMENT;
  if (rate_parameters && parameter_stride < num_cells) return MKPP_ERROR_INVALID_ARGUMENT;
  if (!o->integrate_reduced) return MKPP_ERROR_INVALID_STATE;
  DeviceArgs a{};
  a.conc = conc;
  a.ncells = num_cells;
  a.stride = stride;
  a.temp = temp;
  a.air_density = air_density;
  a.jvals = rate_parameters;
  a.jstride = parameter_stride;
  a.dt = dt;
  a.steps = steps;
  a.method = method;
  a.seed = 1.0;
  a.atol = atol;
  a.rtol = rtol;
  a.status = status;
  a.accepted = accepted;
  a.attempted = attempted;
  a.max_attempts = max_attempts;
  a.importance_threshold = importance_threshold;
  return o->integrate_reduced(a, (hipStream_t)stream) == 0 ? MKPP_SUCCESS : MKPP_ERROR_INVALID_STATE;
}

int mkpp_amd_device_evaluate(const char* mechanism, int what, const double* conc, long num_cells, long stride,
                             const double* temp, const double* air_density, const double* rate_parameters,
                             long parameter_stride, double* out, long out_stride, void* stream) {
  const MechOps* o = mkpp::find_mechanism(mechanism);
  if (!mechanism || !conc || !out) return MKPP_ERROR_NULL_POINTER;
  if (!o) return MKPP_ERROR_NOT_FOUND;
  if (num_cells < 0 || stride < num_cells || out_stride < num_cells || (what != 0 && what != 1))
    return MKPP_ERROR_INVALID_ARGUMENT;
  EvalArgs e{};
  e.conc = conc;
  e.ncells = num_cells;
  e.stride = stride;
  e.temp = temp;
  e.air_density = air_density;
  e.jvals = rate_parameters;
  e.jstride = parameter_stride;
  e.out = out;
  e.out_stride = out_stride;
  const int rc = what == 0 ? o->eval_rates(e, (hipStream_t)stream) : o->eval_jacobian(e, (hipStream_t)stream);
  return rc == 0 ? MKPP_SUCCESS : MKPP_ERROR_INVALID_STATE;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Persistent batch (mkpp_amd.h section 4)
// ---------------------------------------------------------------------------
namespace mkpp {
struct BatchEngineInfo {
  const char* name;
  int wave;  // 1: one cell per wavefront, 0: one cell per lane
  int lds_doubles;
};
#include "gen/batch_list.inc"  // kBatchEngines, kBatchLaunch

static int batch_engine_index(const char* name) {
  for (int i = 0; i < (int)(sizeof(kBatchEngines) / sizeof(kBatchEngines[0])); ++i)
    if (std::strcmp(kBatchEngines[i].name, name) == 0) return i;
  return -1;
}
}  // namespace mkpp

int mkpp_amd_batch_capacity(int method, int* per_cu, int* num_cus) {
  if (!per_cu || !num_cus) return MKPP_ERROR_NULL_POINTER;
  if (method < mkpp::ROS2 || method > mkpp::RODAS4) return MKPP_ERROR_INVALID_ARGUMENT;
  size_t lds = 0;
  for (const auto& e : mkpp::kBatchEngines) lds = std::max(lds, sizeof(double) * (size_t)e.lds_doubles);
  int dev = 0, cus = 0;
  if (hipGetDevice(&dev) != hipSuccess ||
      hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
    return MKPP_ERROR_INVALID_STATE;
  *per_cu = mkpp::kBatchLaunch[method]()->resident_per_cu(lds);
  *num_cus = cus;
  return *per_cu > 0 ? MKPP_SUCCESS : MKPP_ERROR_INVALID_STATE;
}

namespace {
// Both batch entry points: `heads` is device memory for the BATCH_QUEUES
// queue heads (the caller's workspace, or the library's static scratch).
int batch_integrate(const mkpp_amd_batch_segment* segments, int num_segments, double dt, int steps, int method,
                    double initial_step_seed, long max_attempts, int cells_per_lane, int workgroups,
                    unsigned int* heads, void* stream) {
  if (!segments || !heads) return MKPP_ERROR_NULL_POINTER;
  if (num_segments < 1 || num_segments > mkpp::BATCH_MAX_SEGMENTS || !std::isfinite(dt) || dt <= 0.0 || steps < 1 ||
      method < mkpp::ROS2 || method > mkpp::RODAS4 || !(initial_step_seed > 0.0) || cells_per_lane < 1 ||
      cells_per_lane > 64 || workgroups < 0)
    return MKPP_ERROR_INVALID_ARGUMENT;
  mkpp::BatchArgs b{};
  size_t lds = 0;
  // wave-per-cell segments first: their items are the longest, so they start
  // before the cheap lane items that fill the tail
  int order[mkpp::BATCH_MAX_SEGMENTS], n = 0;
  for (int pass = 1; pass >= 0; --pass)
    for (int i = 0; i < num_segments; ++i) {
      const mkpp_amd_batch_segment& g = segments[i];
      if (!g.mechanism || !g.conc || !g.atol || !g.rtol) return MKPP_ERROR_NULL_POINTER;
      const int e = mkpp::batch_engine_index(g.mechanism);
      if (e < 0) return MKPP_ERROR_NOT_FOUND;
      if (g.num_cells < 0 || g.stride < g.num_cells || (g.rate_parameters && g.parameter_stride < g.num_cells))
        return MKPP_ERROR_INVALID_ARGUMENT;
      if (mkpp::kBatchEngines[e].wave == pass) order[n++] = i;
    }
  long per_queue_max = 0;
  for (int k = 0; k < num_segments; ++k) {
    const mkpp_amd_batch_segment& g = segments[order[k]];
    const int e = mkpp::batch_engine_index(g.mechanism);
    mkpp::BatchSegment& s = b.seg[k];
    DeviceArgs& a = s.args;
    a.conc = g.conc;
    a.ncells = g.num_cells;
    a.stride = g.stride;
    a.temp = g.temp;
    a.air_density = g.air_density;
    a.jvals = g.rate_parameters;
    a.jstride = g.parameter_stride;
    a.dt = dt;
    a.steps = steps;
    a.method = method;
    a.seed = initial_step_seed;
    a.atol = g.atol;
    a.rtol = g.rtol;
    a.status = g.status;
    a.accepted = g.accepted;
    a.attempted = g.attempted;
    a.max_attempts = max_attempts;
    s.engine = e;
    if (mkpp::kBatchEngines[e].wave) {
      s.lanes = 0;
      s.items = g.num_cells;
      lds = std::max(lds, sizeof(double) * (size_t)mkpp::kBatchEngines[e].lds_doubles);
    } else {
      s.lanes = cells_per_lane;
      const long per_item = 64L * cells_per_lane;
      s.items = (g.num_cells + per_item - 1) / per_item;
    }
    per_queue_max += s.items / mkpp::BATCH_QUEUES + 1;
  }
  if (per_queue_max >= (1L << 31)) return MKPP_ERROR_INVALID_ARGUMENT;  // 32-bit queue heads
  b.nseg = num_segments;
  b.heads = heads;
  const mkpp::BatchLaunch* L = mkpp::kBatchLaunch[method]();
  if (workgroups == 0) {
    int dev = 0, cus = 0;
    if (hipGetDevice(&dev) != hipSuccess ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess)
      return MKPP_ERROR_INVALID_STATE;
    const int per_cu = L->resident_per_cu(lds);
    if (per_cu <= 0) return MKPP_ERROR_INVALID_STATE;
    workgroups = per_cu * cus;
  }
  return L->launch(b, lds, workgroups, (hipStream_t)stream) == 0 ? MKPP_SUCCESS : MKPP_ERROR_INVALID_STATE;
}
}  // namespace

// Scratch of mkpp_amd_device_integrate_batch_static: the queue heads, in a
// fixed device array of MKPP_AMD_BATCH_STATIC_SCRATCH_BYTES that every GPU the
// library runs on holds until the process ends.  Host code never passes the
// array's name to a kernel (on the host it is only a stand-in for the
// symbol): the device address comes from hipGetSymbolAddress, once per
// device with that device current, and is kept per device.
static __device__ __align__(256) unsigned int scratch[MKPP_AMD_BATCH_STATIC_SCRATCH_BYTES / sizeof(unsigned int)];

namespace {
constexpr int kMaxDevices = 64;
static_assert(sizeof(unsigned int) * mkpp::BATCH_QUEUES <= MKPP_AMD_BATCH_STATIC_SCRATCH_BYTES,
              "the queue heads must fit the static scratch");
static_assert(MKPP_AMD_BATCH_WORKSPACE_BYTES == sizeof(unsigned int) * mkpp::BATCH_QUEUES,
              "the documented workspace size is the queue heads'");
std::mutex g_scratch_mutex;
unsigned int* g_scratch_ptr[kMaxDevices] = {};

// Device address of the current device's scratch (nullptr and *rc set on failure).
unsigned int* static_scratch(int* rc) {
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) {
    *rc = MKPP_ERROR_INVALID_STATE;
    return nullptr;
  }
  std::lock_guard<std::mutex> lock(g_scratch_mutex);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    if (hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch)) != hipSuccess || ptr == nullptr) {
      *rc = MKPP_ERROR_INVALID_STATE;
      return nullptr;
    }
    g_scratch_ptr[dev] = static_cast<unsigned int*>(ptr);
  }
  *rc = MKPP_SUCCESS;
  return g_scratch_ptr[dev];
}
}  // namespace

int mkpp_amd_device_integrate_batch(const mkpp_amd_batch_segment* segments, int num_segments, double dt, int steps,
                                    int method, double initial_step_seed, long max_attempts, int cells_per_lane,
                                    int workgroups, void* workspace, void* stream) {
  if (!workspace) return MKPP_ERROR_NULL_POINTER;
  return batch_integrate(segments, num_segments, dt, steps, method, initial_step_seed, max_attempts, cells_per_lane,
                         workgroups, static_cast<unsigned int*>(workspace), stream);
}

int mkpp_amd_device_integrate_batch_static(const mkpp_amd_batch_segment* segments, int num_segments, double dt,
                                           int steps, int method, double initial_step_seed, long max_attempts,
                                           int cells_per_lane, int workgroups, void* stream) {
  // the heads this launch needs: one 32-bit counter per queue, whatever the input size
  const size_t need = sizeof(unsigned int) * mkpp::BATCH_QUEUES;
  if (need > MKPP_AMD_BATCH_STATIC_SCRATCH_BYTES) return MKPP_ERROR_INVALID_ARGUMENT;
  int rc = MKPP_SUCCESS;
  unsigned int* heads = static_scratch(&rc);
  if (!heads) return rc;
  return batch_integrate(segments, num_segments, dt, steps, method, initial_step_seed, max_attempts, cells_per_lane,
                         workgroups, heads, stream);
}

// ---------------------------------------------------------------------------
// Differentiation (mkpp_amd.h section 5)
// ---------------------------------------------------------------------------
namespace {
const MechOps* diff_mechanism(const char* name, int* rc) {
  const MechOps* o = mkpp::find_mechanism(name);
  if (!name) *rc = MKPP_ERROR_NULL_POINTER;
  else if (!o) *rc = MKPP_ERROR_NOT_FOUND;
  else if (!o->diff) *rc = MKPP_ERROR_INVALID_ARGUMENT;
  else *rc = MKPP_SUCCESS;
  return *rc == MKPP_SUCCESS ? o : nullptr;
}

DeviceArgs diff_args(const double* conc, long num_cells, long stride, const double* temp, const double* air_density,
                     const double* rate_parameters, long parameter_stride) {
  DeviceArgs a{};
  a.conc = const_cast<double*>(conc);
  a.ncells = num_cells;
  a.stride = stride;
  a.temp = temp;
  a.air_density = air_density;
  a.jvals = rate_parameters;
  a.jstride = parameter_stride;
  a.steps = 1;
  return a;
}
}  // namespace

int mkpp_amd_diff_info(const char* mechanism, int* num_groups, int* num_variable) {
  if (!num_groups || !num_variable) return MKPP_ERROR_NULL_POINTER;
  int rc;
  const MechOps* o = diff_mechanism(mechanism, &rc);
  if (!o) return rc;
  *num_groups = o->num_groups;
  *num_variable = o->num_variable;
  return MKPP_SUCCESS;
}

int mkpp_amd_device_rate_constants(const char* mechanism, const double* conc, long num_cells, long stride,
                                   const double* temp, const double* air_density, const double* rate_parameters,
                                   long parameter_stride, double* k, long k_stride, void* stream) {
  int rc;
  const MechOps* o = diff_mechanism(mechanism, &rc);
  if (!o) return rc;
  if (!conc || !k) return MKPP_ERROR_NULL_POINTER;
  if (num_cells < 0 || stride < num_cells || k_stride < num_cells ||
      (rate_parameters && parameter_stride < num_cells))
    return MKPP_ERROR_INVALID_ARGUMENT;
  DeviceArgs a = diff_args(conc, num_cells, stride, temp, air_density, rate_parameters, parameter_stride);
  mkpp::AdjArgs g{};
  g.kbar = k;  // mode 0 writes the rate constants through this pointer
  g.kstride = k_stride;
  return o->diff(0, a, g, (hipStream_t)stream) == 0 ? MKPP_SUCCESS : MKPP_ERROR_INVALID_STATE;
}

int mkpp_amd_device_integrate_checkpointed(const char* mechanism, double* conc, long num_cells, long stride,
                                           const double* temp, const double* air_density,
                                           const double* rate_parameters, long parameter_stride, const double* k,
                                           long k_stride, double dt, int method, double initial_step_seed,
                                           const double* atol, const double* rtol, int* status, int* accepted,
                                           int* attempted, long max_attempts, double* ckpt, double* hs, int* nsteps,
                                           int max_steps, int replay, void* stream) {
  int rc;
  const MechOps* o = diff_mechanism(mechanism, &rc);
  if (!o) return rc;
  if (!conc || !atol || !rtol || !ckpt || !hs || !nsteps) return MKPP_ERROR_NULL_POINTER;
  if (num_cells < 0 || stride < num_cells || (rate_parameters && parameter_stride < num_cells) ||
      (k && k_stride < num_cells) || !std::isfinite(dt) || dt <= 0.0 || !(initial_step_seed > 0.0) ||
      max_steps < 1 || (method < mkpp::ROS2 || method > mkpp::RODAS4) || (replay != 0 && replay != 1))
    return MKPP_ERROR_INVALID_ARGUMENT;
  DeviceArgs a = diff_args(conc, num_cells, stride, temp, air_density, rate_parameters, parameter_stride);
  a.dt = dt;
  a.method = method;
  a.seed = initial_step_seed;
  a.atol = atol;
  a.rtol = rtol;
  a.status = status;
  a.accepted = accepted;
  a.attempted = attempted;
  a.max_attempts = max_attempts;
  mkpp::AdjArgs g{};
  g.k = k;
  g.kstride = k_stride;
  g.ckpt = ckpt;
  g.hs = hs;
  g.nsteps = nsteps;
  g.maxsteps = max_steps;
  return o->diff(replay ? 2 : 1, a, g, (hipStream_t)stream) == 0 ? MKPP_SUCCESS : MKPP_ERROR_INVALID_STATE;
}

int mkpp_amd_device_adjoint(const char* mechanism, const double* conc, long num_cells, long stride,
                            const double* temp, const double* air_density, const double* rate_parameters,
                            long parameter_stride, const double* k, long k_stride, int method, const double* ckpt,
                            const double* hs, const int* nsteps, int max_steps, double* lambda, double* kbar,
                            void* stream) {
  int rc;
  const MechOps* o = diff_mechanism(mechanism, &rc);
  if (!o) return rc;
  if (!conc || !ckpt || !hs || !nsteps || !lambda || !kbar) return MKPP_ERROR_NULL_POINTER;
  if (num_cells < 0 || stride < num_cells || (rate_parameters && parameter_stride < num_cells) ||
      k_stride < num_cells || max_steps < 1 || (method < mkpp::ROS2 || method > mkpp::RODAS4))
    return MKPP_ERROR_INVALID_ARGUMENT;
  DeviceArgs a = diff_args(conc, num_cells, stride, temp, air_density, rate_parameters, parameter_stride);
  a.method = method;
  mkpp::AdjArgs g{};
  g.k = k;
  g.kstride = k_stride;
  g.ckpt = const_cast<double*>(ckpt);
  g.hs = const_cast<double*>(hs);
  g.nsteps = const_cast<int*>(nsteps);
  g.maxsteps = max_steps;
  g.lambda = lambda;
  g.kbar = kbar;
  return o->diff(3, a, g, (hipStream_t)stream) == 0 ? MKPP_SUCCESS : MKPP_ERROR_INVALID_STATE;
}

int mkpp_amd_device_tangent_linear(const char* mechanism, double* conc, long num_cells, long stride,
                                   const double* temp, const double* air_density, const double* rate_parameters,
                                   long parameter_stride, const double* k, long k_stride, int method,
                                   const double* hs, const int* nsteps, int max_steps, double* dy, const double* dk,
                                   int* status, void* stream) {
  int rc;
  const MechOps* o = diff_mechanism(mechanism, &rc);
  if (!o) return rc;
  if (!conc || !hs || !nsteps || !dy) return MKPP_ERROR_NULL_POINTER;
  if (num_cells < 0 || stride < num_cells || (rate_parameters && parameter_stride < num_cells) ||
      ((k || dk) && k_stride < num_cells) || max_steps < 1 || (method < mkpp::ROS2 || method > mkpp::RODAS4))
    return MKPP_ERROR_INVALID_ARGUMENT;
  DeviceArgs a = diff_args(conc, num_cells, stride, temp, air_density, rate_parameters, parameter_stride);
  a.method = method;
  a.status = status;
  mkpp::AdjArgs g{};
  g.k = k;
  g.kstride = k_stride;
  g.hs = const_cast<double*>(hs);
  g.nsteps = const_cast<int*>(nsteps);
  g.maxsteps = max_steps;
  g.dy = dy;
  g.dk = dk;
  return o->diff(4, a, g, (hipStream_t)stream) == 0 ? MKPP_SUCCESS : MKPP_ERROR_INVALID_STATE;
}
