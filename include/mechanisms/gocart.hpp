// gocart.hpp -- SolverKernels of the compiled mechanism gocart.
// Generated by mkpp_amd.gen.emit_host; do not edit.
//
// The reference's generated mechanism header (src/mkpp/templates/header.j2:65-269)
// declares the Species enumeration and SolverKernels<DeviceType> with the
// mechanism's constants and a per-cell integrate() run inside a Kokkos
// parallel_for.  Here one integrate() call runs every cell of a (cells x
// species) view on the GPU through the C ABI of libmkpp_amd.so, with the
// solver options of this type: the generation-time method and tolerances
// below and the initial_step_seed member (which a caller may override).
#pragma once
#include "mkpp_host/dispatcher.hpp"

namespace mkpp::generated::gocart {

// Species indices into the state (physical order).
enum Species : int {
  OH = 0,
  NO3 = 1,
  H2O2 = 2,
  DMS = 3,
  SO2 = 4,
  SO4 = 5,
  BC1 = 6,
  BC2 = 7,
  OC1 = 8,
  OC2 = 9,
  DUST1 = 10,
  DUST2 = 11,
  DUST3 = 12,
  DUST4 = 13,
  DUST5 = 14,
  SEAS1 = 15,
  SEAS2 = 16,
  SEAS3 = 17,
  SEAS4 = 18,
  SEAS5 = 19,
  HNO3 = 20,
  NH3 = 21,
  NH4a = 22,
  NO3an1 = 23,
  NO3an2 = 24,
  NO3an3 = 25
};

template <class DeviceType = ::mkpp::host::DefaultMemorySpace>
struct SolverKernels {
  using memory_space = DeviceType;
  static constexpr const char* NAME = "gocart";
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = true;
  static constexpr int NUM_RUNTIME_PARAMETERS = 1;
  static constexpr bool HAS_CLOUD_GATED = false;
  static constexpr int NUM_SPECIES = 26;
  static constexpr int NUM_REACTIONS = 27;
  static constexpr int NUM_PHOTOLYSIS = 1;
  // generation-time solver (ros3) and tolerances, physical species order
  static constexpr int METHOD = 1;
  static constexpr double atol[NUM_SPECIES] = {
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001
  };
  static constexpr double rtol[NUM_SPECIES] = {
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06
  };
  double initial_step_seed = 1e-06;

  // Integrates every cell of `state` (extent(0) cells x NUM_SPECIES, cells
  // contiguous for device memory) over params.dt, params.steps times, with
  // the per-cell overrides of `params` (dispatcher.hpp:197-255).
  template <class StateView, class MemorySpace>
  void integrate(StateView state, const ::mkpp::host::BasicHostExecutionParams<MemorySpace>& params) const {
    ::mkpp::host::SolverOptions o;
    o.method = METHOD;
    o.atol = atol;
    o.rtol = rtol;
    o.initial_step_seed = initial_step_seed;
    ::mkpp::host::detail::execute_with(NAME, NUM_SPECIES, state, params, o);
  }
};

}  // namespace mkpp::generated::gocart

namespace mech_gocart {
namespace mkpp = ::mkpp::generated::gocart;
}
