// carbon.hpp -- SolverKernels of the compiled mechanism carbon.
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

namespace mkpp::generated::carbon {

// Species indices into the state (physical order).
enum Species : int {
  CH4 = 0,
  CO = 1,
  PCOfromCH4 = 2,
  PCOfromNMVOC = 3,
  LCH4byOH = 4,
  LCH4byCl = 5,
  LCObyOH = 6,
  FixedOH = 7,
  FixedCl = 8,
  DummyCH4 = 9,
  DummyNMVOC = 10
};

template <class DeviceType = ::mkpp::host::DefaultMemorySpace>
struct SolverKernels {
  using memory_space = DeviceType;
  static constexpr const char* NAME = "carbon";
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = true;
  static constexpr int NUM_RUNTIME_PARAMETERS = 0;
  static constexpr bool HAS_CLOUD_GATED = false;
  static constexpr int NUM_SPECIES = 11;
  static constexpr int NUM_REACTIONS = 5;
  static constexpr int NUM_PHOTOLYSIS = 0;
  // generation-time solver (ros3) and tolerances, physical species order
  static constexpr int METHOD = 1;
  static constexpr double atol[NUM_SPECIES] = {
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001
  };
  static constexpr double rtol[NUM_SPECIES] = {
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06
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

}  // namespace mkpp::generated::carbon

namespace mech_carbon {
namespace mkpp = ::mkpp::generated::carbon;
}
