// ts1.hpp -- SolverKernels of the compiled mechanism ts1.
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

namespace mkpp::generated::ts1 {

// Species indices into the state (physical order).
enum Species : int {
  ALKNIT = 0,
  BZOOH = 1,
  C6H5OOH = 2,
  COF2 = 3,
  O2 = 4,
  COFCL = 5,
  HF = 6,
  F = 7,
  BENZO2 = 8,
  BZOO = 9,
  N2 = 10,
  E90 = 11,
  NH_5 = 12,
  NH_50 = 13,
  ST80_25 = 14,
  PAN = 15,
  MVK = 16,
  MACROOH = 17,
  SOAG0 = 18,
  SOAG1 = 19,
  SOAG2 = 20,
  SOAG3 = 21,
  SOAG4 = 22,
  soa4_a1 = 23,
  soa5_a1 = 24,
  soa5_a2 = 25,
  soa3_a1 = 26,
  soa2_a1 = 27,
  soa1_a1 = 28,
  soa1_a2 = 29,
  soa2_a2 = 30,
  soa3_a2 = 31,
  soa4_a2 = 32,
  ISOPNITB = 33,
  SO3 = 34,
  OCS = 35,
  SO = 36,
  S = 37,
  BCARYO2VBS = 38,
  SF6 = 39,
  sink = 40,
  H = 41,
  MEK = 42,
  MTERP = 43,
  N2O5 = 44,
  HCN = 45,
  SVOC = 46,
  ISOPNO3 = 47,
  RO2 = 48,
  PHENO2 = 49,
  BENZO2VBS = 50,
  IVOC = 51,
  TERPNIT = 52,
  ISOPO2VBS = 53,
  IVOCO2VBS = 54,
  HNO3 = 55,
  ACBZO2 = 56,
  CH3COOOH = 57,
  SO2 = 58,
  MTERPO2VBS = 59,
  CH4 = 60,
  CH3CL = 61,
  CH3CO3 = 62,
  C6H5O2 = 63,
  TERPROD1 = 64,
  HYAC = 65,
  HPALD = 66,
  TOLUO2VBS = 67,
  H2O = 68,
  NO2 = 69,
  EOOH = 70,
  NTERPOOH = 71,
  XYLEO2VBS = 72,
  CCL4 = 73,
  CF2CLBR = 74,
  CF3BR = 75,
  CFC11 = 76,
  CFC113 = 77,
  CFC114 = 78,
  CFC115 = 79,
  CFC12 = 80,
  CH2BR2 = 81,
  CH3BR = 82,
  CH3CCL3 = 83,
  NO = 84,
  BR = 85,
  BRCL = 86,
  BRO = 87,
  BRONO2 = 88,
  CL = 89,
  CL2 = 90,
  CL2O2 = 91,
  CLO = 92,
  CLONO2 = 93,
  HCOOH = 94,
  HBR = 95,
  HOBR = 96,
  HOCL = 97,
  N = 98,
  BIGENE = 99,
  C2H4 = 100,
  C2H5O2 = 101,
  CH3COCHO = 102,
  CH3COCH3 = 103,
  O = 104,
  OCLO = 105,
  O1D = 106,
  PHENO = 107,
  HCFC141B = 108,
  HCFC142B = 109,
  HCFC22 = 110,
  DMS = 111,
  C2H5OH = 112,
  HCL = 113,
  BEPOMUC = 114,
  CHBR3 = 115,
  H2402 = 116,
  CO2 = 117,
  BZALD = 118,
  BENZENE = 119,
  C3H7O2 = 120,
  CH3O2 = 121,
  BCARY = 122,
  BIGALD = 123,
  BIGALD2 = 124,
  BIGALD3 = 125,
  BIGALD4 = 126,
  BIGALK = 127,
  H2O2 = 128,
  C2H5OOH = 129,
  C2H6 = 130,
  C3H8 = 131,
  C3H6 = 132,
  CH2O = 133,
  CH3CN = 134,
  C2H2 = 135,
  CH3OH = 136,
  CH3OOH = 137,
  CRESOL = 138,
  ENEO2 = 139,
  MACRO2 = 140,
  ISOPAO2 = 141,
  MALO2 = 142,
  ISOPBO2 = 143,
  MCO3 = 144,
  MDIALO2 = 145,
  MEKO2 = 146,
  EO2 = 147,
  EO = 148,
  GLYOXAL = 149,
  MPAN = 150,
  NC4CH2OH = 151,
  ISOPOOH = 152,
  GLYALD = 153,
  HO2 = 154,
  HOCH2OO = 155,
  H2 = 156,
  HYDRALD = 157,
  ISOP = 158,
  NTERPO2 = 159,
  TOLO2 = 160,
  TERP2O2 = 161,
  XYLENO2 = 162,
  TERPO2 = 163,
  XYLOLO2 = 164,
  PBZNIT = 165,
  XYLENES = 166,
  PO2 = 167,
  TOLUENE = 168,
  XO2 = 169,
  XOOH = 170,
  TERPROD2 = 171,
  MEKOOH = 172,
  MACR = 173,
  HONITR = 174,
  ISOPNITA = 175,
  ISOPNOOH = 176,
  IEPOX = 177,
  ONITR = 178,
  H2SO4 = 179,
  N2O = 180,
  NO3 = 181,
  OH = 182,
  PHENOOH = 183,
  PHENOL = 184,
  XYLOL = 185,
  ROOH = 186,
  O3 = 187,
  TERPOOH = 188,
  TOLOOH = 189,
  XYLENOOH = 190,
  XYLOLOOH = 191,
  CO = 192,
  CH3COOH = 193,
  CH3CHO = 194,
  BIGALD1 = 195,
  ALKOOH = 196,
  DICARBO2 = 197,
  BENZOOH = 198,
  ALKO2 = 199,
  HO2NO2 = 200,
  C3H7OOH = 201,
  NH3 = 202,
  TERP2OOH = 203,
  POOH = 204,
  NOA = 205,
  NC4CHO = 206,
  TEPOMUC = 207,
  NH4 = 208,
  M = 209
};

template <class DeviceType = ::mkpp::host::DefaultMemorySpace>
struct SolverKernels {
  using memory_space = DeviceType;
  static constexpr const char* NAME = "ts1";
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = true;
  static constexpr int NUM_RUNTIME_PARAMETERS = 547;
  static constexpr bool HAS_CLOUD_GATED = false;
  static constexpr int NUM_SPECIES = 210;
  static constexpr int NUM_REACTIONS = 547;
  static constexpr int NUM_PHOTOLYSIS = 547;
  // generation-time solver (ros3) and tolerances, physical species order
  static constexpr int METHOD = 1;
  static constexpr double atol[NUM_SPECIES] = {
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001,
      0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001, 0.001
  };
  static constexpr double rtol[NUM_SPECIES] = {
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06,
      1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06, 1e-06
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

}  // namespace mkpp::generated::ts1

namespace mech_ts1 {
namespace mkpp = ::mkpp::generated::ts1;
}
