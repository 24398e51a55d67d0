// Generated by mkpp_amd.gen.emit_wave: wave-per-cell tables for chapman.
// Do not edit: regenerate with `python -m mkpp_amd.gen`.
#pragma once
#include <cstdint>

namespace mkpp { namespace gen { namespace chapman_wave {

__device__ const int16_t R[3] = {0, 3, 6};
__device__ const uint8_t SLOT[3] = {0, 3, 6};
__device__ const uint8_t ROWOF[64] = {0, 255, 255, 1, 255, 255, 2, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255};
__device__ const uint8_t LCNT[3] = {2, 1, 0};
__device__ const uint16_t LOFS[4] = {0, 2, 3, 3};
__device__ const uint8_t LROW[3] = {1, 2, 2};
__device__ const uint16_t UOFS[4] = {0, 2, 3, 3};
__device__ const uint8_t UCOL[3] = {1, 2, 2};
__device__ const uint64_t FMASK[3] = {6ull, 4ull, 0ull};
__device__ const uint64_t BMASK[3] = {0ull, 1ull, 3ull};
__device__ const uint8_t LAW[4] = {0, 1, 0, 1};
__device__ const int16_t PHOTO[4] = {0, -1, 1, -1};
__device__ const double PARAMS[28] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 6e-34, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 8e-12, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
__device__ const uint8_t GFAC[12] = {2, 4, 4, 1, 2, 3, 0, 4, 4, 1, 0, 4};
__device__ const uint8_t GNPOW[4] = {1, 3, 1, 2};
__device__ const uint8_t GPOWS[12] = {2, 4, 4, 1, 2, 3, 0, 4, 4, 1, 0, 4};
__device__ const uint8_t GPOWP[12] = {1, 0, 0, 1, 1, 1, 1, 0, 0, 1, 1, 0};
__device__ const uint8_t GPOWK[12] = {2, 4, 4, 1, 2, 3, 0, 4, 4, 1, 0, 4};
__device__ const uint16_t POFF[5] = {0, 1, 4, 5, 7};
__device__ const double COEF[4] = {2.0, -1.0, 1.0, 0.0};
__device__ const uint16_t FOFS[4] = {0, 3, 7, 11};
__device__ const uint16_t FG[11] = {1, 2, 3, 0, 1, 2, 3, 0, 1, 2, 3};
__device__ const uint8_t FC[11] = {2, 1, 1, 0, 1, 2, 1, 1, 1, 2, 0};
__device__ const uint16_t WADDR[9] = {0, 1, 3, 4, 5, 6, 7, 8, 2};
__device__ const uint8_t WDIAG[9] = {1, 0, 0, 1, 0, 0, 0, 1, 0};
__device__ const uint16_t WOFS[10] = {0, 2, 4, 6, 8, 10, 12, 14, 16, 17};
__device__ const uint16_t WPI[17] = {4, 6, 1, 5, 4, 6, 1, 5, 0, 2, 4, 6, 1, 5, 0, 2, 2};
__device__ const uint8_t WCI[17] = {1, 1, 2, 1, 2, 1, 1, 1, 0, 1, 2, 0, 1, 0, 1, 1, 2};
__device__ const uint8_t VARPHYS[3] = {2, 0, 1};
__device__ const uint8_t FIXPHYS[1] = {3};
__device__ const uint8_t UNIT[4] = {0, 0, 0, 0};
__device__ const uint8_t GPOWS_PHYS[12] = {1, 0, 0, 0, 1, 3, 2, 0, 0, 0, 2, 0};
__device__ const uint16_t GNUOFS[5] = {0, 2, 5, 8, 11};
__device__ const uint8_t GNUS[11] = {0, 1, 0, 1, 2, 0, 1, 2, 0, 1, 2};
__device__ const uint8_t GNUC[11] = {0, 1, 1, 1, 2, 2, 2, 1, 1, 0, 1};
__device__ const uint16_t GDEPOFS[5] = {0, 1, 4, 5, 7};
__device__ const uint8_t GDEP[7] = {1, 0, 1, 3, 2, 0, 2};
__device__ const uint16_t GNVOFS[5] = {0, 2, 5, 8, 11};
__device__ const uint8_t GNVR[11] = {1, 2, 1, 2, 0, 1, 2, 0, 1, 2, 0};
__device__ const uint8_t GNVC[11] = {0, 1, 1, 1, 2, 2, 2, 1, 1, 0, 1};
__device__ const uint16_t SPOFS[5] = {0, 2, 4, 6, 7};
__device__ const uint16_t SPI[7] = {4, 6, 1, 5, 0, 2, 3};
__device__ const uint16_t MW_ADDR[64] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
__device__ const uint8_t MW_RSLOT[64] = {0, 0, 0, 3, 3, 3, 6, 6, 6, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
__device__ const uint8_t MW_CSLOT[64] = {0, 3, 6, 0, 3, 6, 0, 3, 6, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
__device__ const uint8_t TOLIDX[3] = {1, 2, 3};

struct Wave {
  static constexpr const char* NAME = "chapman";
  static constexpr int N = 4, NV = 3, NF = 1, NG = 4;
  static constexpr int NPH = 2, NB = 1, SPAN = 9, NY = 5;
  static constexpr int NPART = 7, ONE = 4, NENTRIES = 9;
  static constexpr int NFAC = 3;  // factor slots per flux group (GFAC)
  // wavefronts per SIMD the kernels are built for: 2 where the LDS block lets
  // 8 share a CU, else 1 (then spills go to AGPRs, not scratch)
  static constexpr int WAVES_PER_EU = 2;
  static constexpr int NMWP = 1;  // reduction variant: passes over the W locations
  static constexpr int THIRD_BODY = 3, NUM_REACTIONS = 4;
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = false;
  static constexpr int NSCR = 65;  // fluxes or partials + a 0.0 slot; with the row sums: F chunk partials
  static constexpr int LDS_DOUBLES = (SPAN + 1) + NY + NSCR + (NV + 1);
  static constexpr const char* SPECIES[N] = {"O", "O2", "O3", "M"};
};

} } }  // namespace mkpp::gen::chapman_wave
