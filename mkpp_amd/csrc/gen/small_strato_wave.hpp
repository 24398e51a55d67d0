// Generated by mkpp_amd.gen.emit_wave: wave-per-cell tables for small_strato.
// Do not edit: regenerate with `python -m mkpp_amd.gen`.
#pragma once
#include <cstdint>

namespace mkpp { namespace gen { namespace small_strato_wave {

__device__ const int16_t R[6] = {28, 12, 0, 6, 17, 22};
__device__ const uint8_t SLOT[6] = {28, 12, 0, 6, 17, 22};
__device__ const uint8_t ROWOF[64] = {2, 255, 255, 255, 255, 255, 3, 255, 255, 255, 255, 255, 1, 255, 255, 255, 255, 4, 255, 255, 255, 255, 5, 255, 255, 255, 255, 255, 0, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255, 255};
__device__ const uint8_t LCNT[6] = {3, 4, 3, 2, 1, 0};
__device__ const uint16_t LOFS[7] = {0, 3, 7, 10, 12, 13, 13};
__device__ const uint8_t LROW[13] = {1, 2, 3, 2, 3, 4, 5, 3, 4, 5, 4, 5, 5};
__device__ const uint16_t UOFS[7] = {0, 1, 4, 7, 9, 10, 10};
__device__ const uint8_t UCOL[10] = {3, 2, 3, 5, 3, 4, 5, 4, 5, 5};
__device__ const uint64_t FMASK[6] = {14ull, 60ull, 56ull, 48ull, 32ull, 0ull};
__device__ const uint64_t BMASK[6] = {0ull, 0ull, 2ull, 7ull, 12ull, 30ull};
__device__ const uint8_t LAW[10] = {0, 1, 0, 1, 0, 1, 1, 1, 1, 0};
__device__ const int16_t PHOTO[10] = {0, -1, 1, -1, 2, -1, -1, -1, -1, 3};
__device__ const double PARAMS[70] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 8.018e-17, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.576e-15, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 7.11e-11, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.2e-10, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 6.062e-15, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 1.069e-11, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
__device__ const uint8_t GFAC[30] = {2, 7, 7, 1, 2, 7, 3, 7, 7, 1, 3, 7, 3, 7, 7, 0, 6, 7, 0, 3, 7, 3, 4, 7, 1, 5, 7, 5, 7, 7};
__device__ const uint8_t GNPOW[10] = {1, 2, 1, 2, 1, 2, 2, 2, 2, 1};
__device__ const uint8_t GPOWS[30] = {2, 7, 7, 1, 2, 7, 3, 7, 7, 1, 3, 7, 3, 7, 7, 0, 6, 7, 0, 3, 7, 3, 4, 7, 1, 5, 7, 5, 7, 7};
__device__ const uint8_t GPOWP[30] = {1, 0, 0, 1, 1, 0, 1, 0, 0, 1, 1, 0, 1, 0, 0, 1, 1, 0, 1, 1, 0, 1, 1, 0, 1, 1, 0, 1, 0, 0};
__device__ const uint8_t GPOWK[30] = {2, 7, 7, 1, 2, 7, 3, 7, 7, 1, 3, 7, 3, 7, 7, 0, 6, 7, 0, 3, 7, 3, 4, 7, 1, 5, 7, 5, 7, 7};
__device__ const uint16_t POFF[11] = {0, 1, 3, 4, 6, 7, 9, 11, 13, 15, 16};
__device__ const double COEF[4] = {2.0, -1.0, 1.0, 0.0};
__device__ const uint16_t FOFS[7] = {0, 3, 10, 18, 24, 27, 30};
__device__ const uint16_t FG[30] = {4, 5, 6, 0, 1, 2, 3, 5, 8, 9, 0, 1, 2, 3, 4, 6, 7, 8, 1, 2, 3, 4, 6, 7, 7, 8, 9, 7, 8, 9};
__device__ const uint8_t FC[30] = {2, 1, 1, 0, 1, 2, 1, 2, 1, 2, 1, 1, 2, 0, 2, 0, 2, 2, 2, 1, 1, 1, 1, 1, 1, 2, 2, 2, 1, 1};
__device__ const uint16_t WADDR[29] = {3, 9, 13, 1, 28, 31, 14, 15, 17, 2, 7, 22, 27, 12, 0, 4, 5, 6, 8, 10, 18, 20, 21, 23, 25, 26, 11, 19, 24};
__device__ const uint8_t WDIAG[29] = {0, 1, 1, 0, 1, 0, 0, 0, 0, 1, 0, 0, 1, 0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0, 0, 0, 0};
__device__ const uint16_t WOFS[30] = {0, 5, 10, 13, 16, 18, 20, 22, 24, 26, 28, 30, 32, 34, 35, 36, 37, 38, 39, 40, 41, 42, 43, 44, 45, 46, 47, 47, 47, 47};
__device__ const uint16_t WPI[47] = {3, 5, 6, 10, 11, 3, 5, 6, 10, 11, 1, 4, 13, 1, 4, 13, 7, 9, 6, 10, 0, 2, 3, 5, 14, 15, 0, 2, 1, 4, 14, 15, 14, 15, 7, 9, 12, 14, 9, 2, 12, 13, 11, 12, 13, 11, 12};
__device__ const uint8_t WCI[47] = {2, 0, 2, 0, 2, 1, 1, 1, 1, 1, 1, 1, 1, 1, 0, 2, 1, 1, 2, 1, 0, 1, 2, 1, 1, 2, 1, 1, 2, 1, 2, 2, 1, 1, 2, 0, 2, 2, 1, 2, 1, 2, 1, 1, 1, 2, 2};
__device__ const uint8_t VARPHYS[6] = {1, 0, 6, 2, 3, 4};
__device__ const uint8_t FIXPHYS[1] = {5};
__device__ const uint8_t UNIT[7] = {0, 0, 0, 0, 0, 0, 0};
__device__ const uint8_t GPOWS_PHYS[30] = {6, 0, 0, 0, 6, 0, 2, 0, 0, 0, 2, 0, 2, 0, 0, 1, 5, 0, 1, 2, 0, 2, 3, 0, 0, 4, 0, 4, 0, 0};
__device__ const uint16_t GNUOFS[11] = {0, 2, 5, 8, 11, 14, 16, 19, 23, 27, 30};
__device__ const uint8_t GNUS[30] = {0, 6, 0, 2, 6, 0, 2, 6, 0, 2, 6, 1, 2, 6, 0, 1, 1, 2, 6, 2, 3, 4, 6, 0, 3, 4, 6, 0, 3, 4};
__device__ const uint8_t GNUC[30] = {0, 1, 1, 2, 1, 2, 1, 2, 1, 1, 0, 2, 1, 2, 2, 1, 1, 1, 0, 1, 1, 2, 2, 1, 2, 1, 2, 2, 2, 1};
__device__ const uint16_t GDEPOFS[11] = {0, 1, 3, 4, 6, 7, 9, 11, 13, 15, 16};
__device__ const uint8_t GDEP[16] = {6, 0, 6, 2, 0, 2, 2, 1, 5, 1, 2, 2, 3, 0, 4, 4};
__device__ const uint16_t GNVOFS[11] = {0, 2, 5, 8, 11, 14, 16, 19, 23, 27, 30};
__device__ const uint8_t GNVR[30] = {1, 2, 1, 3, 2, 1, 3, 2, 1, 3, 2, 0, 3, 2, 1, 0, 0, 3, 2, 3, 4, 5, 2, 1, 4, 5, 2, 1, 4, 5};
__device__ const uint8_t GNVC[30] = {0, 1, 1, 2, 1, 2, 1, 2, 1, 1, 0, 2, 1, 2, 2, 1, 1, 1, 0, 1, 1, 2, 2, 1, 2, 1, 2, 2, 2, 1};
__device__ const uint16_t SPOFS[8] = {0, 2, 5, 7, 12, 13, 15, 16};
__device__ const uint16_t SPI[16] = {7, 9, 1, 4, 13, 0, 2, 3, 5, 6, 10, 11, 12, 14, 15, 8};
__device__ const uint16_t MW_ADDR[64] = {28, 31, 12, 13, 14, 15, 17, 0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 18, 20, 21, 22, 23, 25, 26, 27, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28};
__device__ const uint8_t MW_RSLOT[64] = {28, 28, 12, 12, 12, 12, 12, 0, 0, 0, 0, 0, 0, 6, 6, 6, 6, 6, 17, 17, 17, 17, 22, 22, 22, 22, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28};
__device__ const uint8_t MW_CSLOT[64] = {28, 6, 28, 12, 0, 6, 22, 28, 12, 0, 6, 17, 22, 28, 12, 0, 6, 17, 12, 6, 17, 22, 12, 6, 17, 22, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28, 28};
__device__ const uint8_t TOLIDX[6] = {1, 2, 3, 4, 5, 6};

struct Wave {
  static constexpr const char* NAME = "small_strato";
  static constexpr int N = 7, NV = 6, NF = 1, NG = 10;
  static constexpr int NPH = 4, NB = 1, SPAN = 32, NY = 8;
  static constexpr int NPART = 16, ONE = 7, NENTRIES = 29;
  static constexpr int NFAC = 3;  // factor slots per flux group (GFAC)
  // wavefronts per SIMD the kernels are built for: 2 where the LDS block lets
  // 8 share a CU, else 1 (then spills go to AGPRs, not scratch)
  static constexpr int WAVES_PER_EU = 2;
  static constexpr int NMWP = 1;  // reduction variant: passes over the W locations
  static constexpr int THIRD_BODY = 5, NUM_REACTIONS = 10;
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = false;
  static constexpr int NSCR = 68;  // fluxes or partials + a 0.0 slot; with the row sums: F chunk partials
  static constexpr int LDS_DOUBLES = (SPAN + 1) + NY + NSCR + (NV + 1);
  static constexpr const char* SPECIES[N] = {"O", "O1D", "O3", "NO", "NO2", "M", "O2"};
};

} } }  // namespace mkpp::gen::small_strato_wave
