// saprcnov.hpp -- SolverKernels of the compiled mechanism saprcnov.
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

namespace mkpp::generated::saprcnov {

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
  ETOH = 19,
  CCO_OH = 20,
  RCO_OH = 21,
  GLY = 22,
  MGLY = 23,
  BACL = 24,
  CRES = 25,
  BALD = 26,
  ISOPROD = 27,
  METHACRO = 28,
  MVK = 29,
  PROD2 = 30,
  DCB1 = 31,
  DCB2 = 32,
  DCB3 = 33,
  ETHENE = 34,
  ISOPRENE = 35,
  C2H6 = 36,
  C3H8 = 37,
  C2H2 = 38,
  C3H6 = 39,
  ALK3 = 40,
  ALK4 = 41,
  ALK5 = 42,
  ARO1 = 43,
  ARO2 = 44,
  OLE1 = 45,
  OLE2 = 46,
  TERP = 47,
  RNO3 = 48,
  NPHE = 49,
  PHEN = 50,
  PAN = 51,
  PAN2 = 52,
  PBZN = 53,
  MA_PAN = 54,
  BC = 55,
  OC = 56,
  SSF = 57,
  SSC = 58,
  PM10 = 59,
  PM25 = 60,
  DMS = 61,
  DST1 = 62,
  DST2 = 63,
  DST3 = 64,
  CO2 = 65,
  CCO_OOH = 66,
  RCO_O2 = 67,
  RCO_OOH = 68,
  XN = 69,
  XC = 70,
  O3P = 71,
  O1D = 72,
  OH = 73,
  HO2 = 74,
  C_O2 = 75,
  COOH = 76,
  ROOH = 77,
  RO2_R = 78,
  R2O2 = 79,
  RO2_N = 80,
  HOCOO = 81,
  CCO_O2 = 82,
  BZCO_O2 = 83,
  BZNO2_O = 84,
  BZ_O = 85,
  MA_RCO3 = 86,
  TBU_O = 87,
  AIR = 88,
  N2 = 89,
  O2 = 90,
  H2O = 91,
  H2 = 92,
  CH4 = 93,
  M = 94,
  RO2 = 95
};

template <class DeviceType = ::mkpp::host::DefaultMemorySpace>
struct SolverKernels {
  using memory_space = DeviceType;
  static constexpr const char* NAME = "saprcnov";
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = true;
  static constexpr int NUM_RUNTIME_PARAMETERS = 29;
  static constexpr bool HAS_CLOUD_GATED = false;
  static constexpr int NUM_SPECIES = 96;
  static constexpr int NUM_REACTIONS = 187;
  static constexpr int NUM_PHOTOLYSIS = 29;
  // generation-time solver (ros3) and tolerances, physical species order
  static constexpr int METHOD = 1;
  static constexpr double atol[NUM_SPECIES] = {
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
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

}  // namespace mkpp::generated::saprcnov

namespace mech_saprcnov {
namespace mkpp = ::mkpp::generated::saprcnov;
}
