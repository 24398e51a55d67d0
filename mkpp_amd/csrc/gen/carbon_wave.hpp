// Generated by mkpp_amd.gen.emit_wave: wave-per-cell tables for carbon.
// Do not edit: regenerate with `python -m mkpp_amd.gen`.
#pragma once
#include <cstdint>

namespace mkpp { namespace gen { namespace carbon_wave {

__device__ const int16_t R[11] = {11, 18, 3, 24, 0, 3, 6, 9, 12, 15, 18};
__device__ const uint8_t SLOT[11] = {11, 18, 35, 24, 0, 3, 6, 9, 12, 15, 50};
__device__ const uint8_t ROWOF[64] = {4, 255, 255, 5, 255, 255, 6, 255, 255, 7, 255, 0, 8, 255, 255, 9, 255, 255, 1, 255, 255, 255, 255, 255, 3, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 2, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 10, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255};
__device__ const uint8_t LCNT[11] = {0, 0, 1, 1, 2, 0, 2, 0, 2, 0, 0};
__device__ const uint16_t LOFS[12] = {0, 0, 0, 1, 2, 4, 4, 6, 6, 8, 8, 8};
__device__ const uint8_t LROW[8] = {4, 4, 5, 6, 7, 8, 9, 10};
__device__ const uint16_t UOFS[12] = {0, 1, 2, 2, 2, 3, 4, 5, 6, 7, 8, 8};
__device__ const uint8_t UCOL[8] = {2, 3, 6, 6, 8, 8, 10, 10};
__device__ const uint64_t FMASK[11] = {0ull, 0ull, 16ull, 16ull, 96ull, 0ull, 384ull, 0ull, 1536ull, 0ull, 0ull};
__device__ const uint64_t BMASK[11] = {0ull, 0ull, 1ull, 2ull, 0ull, 0ull, 48ull, 0ull, 192ull, 0ull, 768ull};
__device__ const uint8_t LAW[5] = {1, 1, 1, 1, 1};
__device__ const int16_t PHOTO[5] = {-1, -1, -1, -1, -1};
__device__ const double PARAMS[35] = {2.45e-12, 0.0, -1775.0, 0.0, 0.0, 0.0, 0.0, 9.6e-12, 0.0, -1360.0, 0.0, 0.0, 0.0, 0.0, 4.2566446e-15, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 7.3679649e-14, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 38199.012, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
__device__ const uint8_t GFAC[15] = {8, 11, 11, 8, 11, 11, 11, 11, 11, 4, 11, 11, 11, 11, 11};
__device__ const uint8_t GNPOW[5] = {2, 2, 1, 2, 1};
__device__ const uint8_t GPOWS[15] = {8, 11, 11, 8, 11, 11, 11, 11, 11, 4, 11, 11, 11, 11, 11};
__device__ const uint8_t GPOWP[15] = {1, 1, 0, 1, 1, 0, 1, 0, 0, 1, 1, 0, 1, 0, 0};
__device__ const uint8_t GPOWK[15] = {8, 6, 11, 8, 10, 11, 3, 11, 11, 4, 6, 11, 2, 11, 11};
__device__ const uint16_t POFF[6] = {0, 2, 4, 5, 7, 8};
__device__ const double COEF[3] = {-1.0, 1.0, 0.0};
__device__ const uint16_t FOFS[12] = {0, 1, 2, 3, 4, 7, 8, 10, 11, 13, 14, 15};
__device__ const uint16_t FG[15] = {4, 2, 4, 2, 2, 3, 4, 3, 0, 3, 0, 0, 1, 1, 1};
__device__ const uint8_t FC[15] = {1, 1, 0, 0, 1, 0, 1, 1, 0, 0, 1, 0, 0, 1, 0};
__device__ const uint16_t WADDR[27] = {12, 20, 13, 21, 5, 27, 2, 3, 4, 6, 7, 9, 10, 14, 15, 17, 18, 22, 23, 25, 26, 28, 11, 19, 8, 16, 24};
__device__ const uint8_t WDIAG[27] = {1, 1, 0, 0, 1, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1, 1, 1};
__device__ const uint16_t WOFS[28] = {0, 2, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 24, 24, 24, 24, 24};
__device__ const uint16_t WPI[24] = {1, 6, 0, 2, 7, 4, 7, 4, 7, 4, 5, 6, 5, 6, 5, 0, 1, 0, 1, 3, 2, 3, 2, 3};
__device__ const uint8_t WCI[24] = {0, 0, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0, 1, 1, 0, 0};
__device__ const uint8_t VARPHYS[11] = {3, 2, 10, 9, 1, 6, 7, 4, 0, 5, 8};
__device__ const uint8_t FIXPHYS[1] = {0};
__device__ const uint8_t UNIT[11] = {0, 0, 0, 0, 0, 0, 0, 1, 1, 1, 1};
__device__ const uint8_t GPOWS_PHYS[15] = {0, 7, 0, 0, 8, 0, 9, 0, 0, 1, 7, 0, 10, 0, 0};
__device__ const uint16_t GNUOFS[6] = {0, 3, 6, 9, 12, 15};
__device__ const uint8_t GNUS[15] = {0, 4, 7, 0, 5, 8, 1, 2, 9, 1, 6, 7, 1, 3, 10};
__device__ const uint8_t GNUC[15] = {0, 1, 0, 0, 1, 0, 1, 1, 0, 0, 1, 0, 1, 1, 0};
__device__ const uint16_t GDEPOFS[6] = {0, 2, 4, 5, 7, 8};
__device__ const uint8_t GDEP[8] = {0, 7, 0, 8, 9, 1, 7, 10};
__device__ const uint16_t GNVOFS[6] = {0, 3, 6, 9, 12, 15};
__device__ const uint8_t GNVR[15] = {8, 7, 6, 8, 9, 10, 4, 1, 3, 4, 5, 6, 4, 0, 2};
__device__ const uint8_t GNVC[15] = {0, 1, 0, 0, 1, 0, 1, 1, 0, 0, 1, 0, 1, 1, 0};
__device__ const uint16_t SPOFS[12] = {0, 0, 0, 0, 0, 1, 1, 1, 1, 3, 3, 3};
__device__ const uint16_t SPI[3] = {5, 0, 2};
__device__ const uint16_t MW_ADDR[64] = {11, 13, 19, 21, 5, 27, 2, 3, 4, 6, 7, 8, 9, 10, 12, 14, 15, 16, 17, 18, 20, 22, 23, 24, 25, 26, 28, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11};
__device__ const uint8_t MW_RSLOT[64] = {11, 11, 18, 18, 35, 24, 0, 0, 0, 0, 3, 3, 3, 6, 6, 6, 9, 9, 9, 12, 12, 12, 15, 15, 15, 50, 50, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11};
__device__ const uint8_t MW_CSLOT[64] = {11, 35, 18, 24, 35, 24, 35, 24, 0, 6, 0, 3, 6, 0, 6, 12, 6, 9, 12, 6, 12, 50, 12, 15, 50, 12, 50, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11, 11};
__device__ const uint8_t TOLIDX[11] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10};

struct Wave {
  static constexpr const char* NAME = "carbon";
  static constexpr int N = 11, NV = 11, NF = 0, NG = 5;
  static constexpr int NPH = 0, NB = 1, SPAN = 29, NY = 12;
  static constexpr int NPART = 8, ONE = 11, NENTRIES = 27;
  static constexpr int NFAC = 3;  // factor slots per flux group (GFAC)
  // wavefronts per SIMD the kernels are built for: 2 where the LDS block lets
  // 8 share a CU, else 1 (then spills go to AGPRs, not scratch)
  static constexpr int WAVES_PER_EU = 2;
  static constexpr int NMWP = 1;  // reduction variant: passes over the W locations
  static constexpr int THIRD_BODY = -1, NUM_REACTIONS = 5;
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = true;
  static constexpr int NSCR = 58;  // fluxes or partials + a 0.0 slot; with the row sums: F chunk partials
  static constexpr int LDS_DOUBLES = (SPAN + 1) + NY + NSCR + (NV + 1);
  static constexpr const char* SPECIES[N] = {"CH4", "CO", "PCOfromCH4", "PCOfromNMVOC", "LCH4byOH", "LCH4byCl", "LCObyOH", "FixedOH", "FixedCl", "DummyCH4", "DummyNMVOC"};
};

} } }  // namespace mkpp::gen::carbon_wave
