// Generated by mkpp_amd.gen.emit_wave: wave-per-cell tables for decay.
// Do not edit: regenerate with `python -m mkpp_amd.gen`.
#pragma once
#include <cstdint>

namespace mkpp { namespace gen { namespace decay_wave {

__device__ const int16_t R[1] = {0};
__device__ const uint8_t SLOT[1] = {0};
__device__ const uint8_t ROWOF[64] = {0, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255};
__device__ const uint8_t LCNT[1] = {0};
__device__ const uint16_t LOFS[2] = {0, 0};
__device__ const uint8_t LROW[1] = {0};
__device__ const uint16_t UOFS[2] = {0, 0};
__device__ const uint8_t UCOL[1] = {0};
__device__ const uint64_t FMASK[1] = {0ull};
__device__ const uint64_t BMASK[1] = {0ull};
__device__ const uint8_t LAW[1] = {0};
__device__ const int16_t PHOTO[1] = {0};
__device__ const double PARAMS[7] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
__device__ const uint8_t GFAC[3] = {0, 1, 1};
__device__ const uint8_t GNPOW[1] = {1};
__device__ const uint8_t GPOWS[3] = {0, 1, 1};
__device__ const uint8_t GPOWP[3] = {1, 0, 0};
__device__ const uint8_t GPOWK[3] = {0, 1, 1};
__device__ const uint16_t POFF[2] = {0, 1};
__device__ const double COEF[2] = {-1.0, 0.0};
__device__ const uint16_t FOFS[2] = {0, 1};
__device__ const uint16_t FG[1] = {0};
__device__ const uint8_t FC[1] = {0};
__device__ const uint16_t WADDR[1] = {0};
__device__ const uint8_t WDIAG[1] = {1};
__device__ const uint16_t WOFS[2] = {0, 1};
__device__ const uint16_t WPI[1] = {0};
__device__ const uint8_t WCI[1] = {0};
__device__ const uint8_t VARPHYS[1] = {0};
__device__ const uint8_t FIXPHYS[1] = {0};
__device__ const uint8_t UNIT[1] = {0};
__device__ const uint8_t GPOWS_PHYS[3] = {0, 0, 0};
__device__ const uint16_t GNUOFS[2] = {0, 1};
__device__ const uint8_t GNUS[1] = {0};
__device__ const uint8_t GNUC[1] = {0};
__device__ const uint16_t GDEPOFS[2] = {0, 1};
__device__ const uint8_t GDEP[1] = {0};
__device__ const uint16_t GNVOFS[2] = {0, 1};
__device__ const uint8_t GNVR[1] = {0};
__device__ const uint8_t GNVC[1] = {0};
__device__ const uint16_t SPOFS[2] = {0, 1};
__device__ const uint16_t SPI[1] = {0};
__device__ const uint16_t MW_ADDR[64] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
__device__ const uint8_t MW_RSLOT[64] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
__device__ const uint8_t MW_CSLOT[64] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
__device__ const uint8_t TOLIDX[1] = {0};

struct Wave {
  static constexpr const char* NAME = "decay";
  static constexpr int N = 1, NV = 1, NF = 0, NG = 1;
  static constexpr int NPH = 1, NB = 1, SPAN = 1, NY = 2;
  static constexpr int NPART = 1, ONE = 1, NENTRIES = 1;
  static constexpr int NFAC = 3;  // factor slots per flux group (GFAC)
  // wavefronts per SIMD the kernels are built for: 2 where the LDS block lets
  // 8 share a CU, else 1 (then spills go to AGPRs, not scratch)
  static constexpr int WAVES_PER_EU = 2;
  static constexpr int NMWP = 1;  // reduction variant: passes over the W locations
  static constexpr int THIRD_BODY = -1, NUM_REACTIONS = 1;
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = false;
  static constexpr int NSCR = 64;  // fluxes or partials + a 0.0 slot; with the row sums: F chunk partials
  static constexpr int LDS_DOUBLES = (SPAN + 1) + NY + NSCR + (NV + 1);
  static constexpr const char* SPECIES[N] = {"X"};
};

} } }  // namespace mkpp::gen::decay_wave
