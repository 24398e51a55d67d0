// saprc99.hpp -- SolverKernels of the compiled mechanism saprc99.
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

namespace mkpp::generated::saprc99 {

// Species indices into the state (physical order).
enum Species : int {
  O3 = 0,
  H2O2 = 1,
  NO = 2,
  NO2 = 3,
  NO3 = 4,
  N2O5 = 5,
  HONO = 6,
  HNO3 = 7,
  HNO4 = 8,
  SO2 = 9,
  H2SO4 = 10,
  CO = 11,
  HCHO = 12,
  CCHO = 13,
  RCHO = 14,
  ACET = 15,
  MEK = 16,
  HCOOH = 17,
  MEOH = 18,
  CCO_OH = 19,
  RCO_OH = 20,
  GLY = 21,
  MGLY = 22,
  BACL = 23,
  CRES = 24,
  BALD = 25,
  ISOPROD = 26,
  METHACRO = 27,
  MVK = 28,
  PROD2 = 29,
  DCB1 = 30,
  DCB2 = 31,
  DCB3 = 32,
  ETHENE = 33,
  ISOPRENE = 34,
  ALK1 = 35,
  ALK2 = 36,
  ALK3 = 37,
  ALK4 = 38,
  ALK5 = 39,
  ARO1 = 40,
  ARO2 = 41,
  OLE1 = 42,
  OLE2 = 43,
  TERP = 44,
  RNO3 = 45,
  NPHE = 46,
  PHEN = 47,
  PAN = 48,
  PAN2 = 49,
  PBZN = 50,
  MA_PAN = 51,
  CCO_OOH = 52,
  RCO_O2 = 53,
  RCO_OOH = 54,
  XN = 55,
  XC = 56,
  O3P = 57,
  O1D = 58,
  OH = 59,
  HO2 = 60,
  C_O2 = 61,
  COOH = 62,
  ROOH = 63,
  RO2_R = 64,
  R2O2 = 65,
  RO2_N = 66,
  HOCOO = 67,
  CCO_O2 = 68,
  BZCO_O2 = 69,
  BZNO2_O = 70,
  BZ_O = 71,
  MA_RCO3 = 72,
  TBU_O = 73,
  AIR = 74,
  O2 = 75,
  H2O = 76,
  H2 = 77,
  CH4 = 78,
  M = 79,
  N2 = 80,
  RO2 = 81
};

template <class DeviceType = ::mkpp::host::DefaultMemorySpace>
struct SolverKernels {
  using memory_space = DeviceType;
  static constexpr const char* NAME = "saprc99";
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = true;
  static constexpr int NUM_RUNTIME_PARAMETERS = 30;
  static constexpr bool HAS_CLOUD_GATED = false;
  static constexpr int NUM_SPECIES = 82;
  static constexpr int NUM_REACTIONS = 211;
  static constexpr int NUM_PHOTOLYSIS = 30;
  // generation-time solver (ros3) and tolerances, physical species order
  static constexpr int METHOD = 1;
  static constexpr double atol[NUM_SPECIES] = {
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001
  };
  static constexpr double rtol[NUM_SPECIES] = {
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
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

}  // namespace mkpp::generated::saprc99

namespace mech_saprc99 {
namespace mkpp = ::mkpp::generated::saprc99;
}
