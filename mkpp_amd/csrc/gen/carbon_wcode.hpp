// Generated by mkpp_amd.gen.emit_wave: straight-line LU and solves for carbon.
// Do not edit: regenerate with `python -m mkpp_amd.gen`.
#pragma once
#include "carbon_wave.hpp"

namespace mkpp { namespace gen { namespace carbon_wave {

__device__ const uint32_t __attribute__((aligned(16))) CF_TERM[256] = {33882372, 33882629, 33882629, 33882629, 33882370, 33882629, 33882629, 33882629, 33882116, 33882629, 33882629, 33882629, 33882114, 33882629, 33882629, 33882629, 196866, 33882372, 33882629, 33882629, 33882371, 33882629, 33882629, 33882629, 196608, 33882629, 33882629, 33882629, 33882368, 33882629, 33882629, 33882629, 65536, 33882629, 33882629, 33882629, 33882369, 33882629, 33882629, 33882629, 33882113, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629, 33882629};
__device__ const uint16_t __attribute__((aligned(8))) CF_PART[256] = {10, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 11, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 12, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 13, 5, 5, 5, 5, 5, 5, 5, 6, 5, 5, 5, 14, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 15, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 7, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 9, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 8, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 16, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5, 5};
__device__ const uint32_t CW_TERM[128] = {1, 0, 65543, 65540, 7, 4, 65543, 65540, 5, 6, 65541, 65542, 5, 0, 65537, 65536, 1, 3, 65538, 65539, 2, 3, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 6, 2, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080, 131080};
__device__ const uint16_t CW_ADDR[64] = {12, 20, 13, 21, 5, 27, 2, 3, 4, 6, 7, 9, 10, 14, 15, 17, 18, 22, 23, 25, 26, 28, 8, 11, 16, 19, 24, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29, 29};
__device__ const uint32_t __attribute__((aligned(16))) CS_TERM[256] = {524293, 524296, 524296, 524296, 131072, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296, 524296};
__device__ const uint16_t __attribute__((aligned(8))) CS_PART[256] = {0, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 1, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64};
__device__ const uint16_t CS_FIX[1] = {64};
__device__ const uint16_t CG_TERM[256] = {8, 8, 260, 4, 260, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 263, 265, 257, 261, 256, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 6, 10, 3, 6, 2, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523, 523};
__device__ const uint16_t CG_GRP[64] = {0, 1, 2, 3, 4, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535};

struct Code {
  static constexpr int TF = 8, NFP = 1, TW = 2, NWP = 1;
  static constexpr int FBITS = 8;  // flux-slot bits of a 16-bit F term (coefficient above)
  // F partial slots scr[PBASE, PBASE + chunks); a row reads FMAXC[bank] of its FMAXCP-wide list
  static constexpr int PBASE = 6, FSLOTS = 17, FMAXCP = 4;
  static constexpr int FMAXC[1] = {1};
  static constexpr int TS = 8, NSP = 1;  // adjoint scatter chunks
  // scatter chunk partials cp[0, 64 NSP), cp[SZERO] = 0.0; rows read SMAXC[bank] of
  // their SMAXCP-wide list, fixed species SFMAXC of SFMAXCP
  static constexpr int SZERO = 64, SMAXCP = 4, SFMAXC = 1, SFMAXCP = 1;
  static constexpr int SMAXC[1] = {1};
  static constexpr int TG = 4, NGAT = 1;  // adjoint gather chunks

  // Right-looking factorization: pivot k divides column k of every row in
  // L(k) by U(k,k) and subtracts l * U(k, j) for j in U(k), each product in
  // ascending-k order per entry (the reference's per-entry sequence).
  static __device__ __forceinline__ void factor(double* __restrict__ lu, const int (&rb)[1], int lane) {
    // pivot rows are read through an opaque base: with a constant address the
    // compiler hoists every one out of the step loop and spills them
    int zero = 0;
    asm volatile("" : "+v"(zero));
    const double* pv = lu + zero;  // aliases lu: pivot rows are read after their stores
    int zs = 0;  // laundered once per call: see lane_mask
    asm volatile("" : "+s"(zs));
    // Rows and columns 0..10 in registers: lane = row (bank 0), c<j> =
    // the row's entry in column j (0 where the pattern has none).  Pivot rows
    // are broadcast with readlane; the multiplier is zeroed outside L(k), so
    // every update is an unconditional fma that leaves other lanes unchanged.
    double c0 = lane_keep(lane_mask<0x00000800u, 0x00000000u, 1>(zs), lu[rb[0] + 0]);
    double c1 = lane_keep(lane_mask<0x00040000u, 0x00000000u, 2>(zs), lu[rb[0] + 1]);
    double c2 = lane_keep(lane_mask<0x00000801u, 0x00000008u, 3>(zs), lu[rb[0] + 2]);
    double c3 = lane_keep(lane_mask<0x01040001u, 0x00000000u, 4>(zs), lu[rb[0] + 3]);
    double c4 = lane_keep(lane_mask<0x00000049u, 0x00000000u, 5>(zs), lu[rb[0] + 4]);
    double c5 = lane_keep(lane_mask<0x00000008u, 0x00000000u, 6>(zs), lu[rb[0] + 5]);
    double c6 = lane_keep(lane_mask<0x00001249u, 0x00000000u, 7>(zs), lu[rb[0] + 6]);
    double c7 = lane_keep(lane_mask<0x00000200u, 0x00000000u, 8>(zs), lu[rb[0] + 7]);
    double c8 = lane_keep(lane_mask<0x00009240u, 0x00040000u, 9>(zs), lu[rb[0] + 8]);
    double c9 = lane_keep(lane_mask<0x00008000u, 0x00000000u, 10>(zs), lu[rb[0] + 9]);
    double c10 = lane_keep(lane_mask<0x00009000u, 0x00040000u, 11>(zs), lu[rb[0] + 10]);
    {  // pivot 2: 1 rows x 0 columns
      const double d = readlane_d(c2, 35);
      const bool in_l = __builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000001u, 0x00000000u, 12>(zs));
      const double l = in_l ? c2 / d : 0.0;
      c2 = in_l ? l : c2;
    }
    {  // pivot 3: 1 rows x 0 columns
      const double d = readlane_d(c3, 24);
      const bool in_l = __builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000001u, 0x00000000u, 13>(zs));
      const double l = in_l ? c3 / d : 0.0;
      c3 = in_l ? l : c3;
    }
    {  // pivot 4: 2 rows x 1 columns
      const double d = readlane_d(c4, 0);
      const bool in_l = __builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000048u, 0x00000000u, 14>(zs));
      const double l = in_l ? c4 / d : 0.0;
      c4 = in_l ? l : c4;
      const double u6 = readlane_d(c6, 0);
      c6 = fma(-l, u6, c6);
    }
    {  // pivot 6: 2 rows x 1 columns
      const double d = readlane_d(c6, 6);
      const bool in_l = __builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00001200u, 0x00000000u, 15>(zs));
      const double l = in_l ? c6 / d : 0.0;
      c6 = in_l ? l : c6;
      const double u8 = readlane_d(c8, 6);
      c8 = fma(-l, u8, c8);
    }
    {  // pivot 8: 2 rows x 1 columns
      const double d = readlane_d(c8, 12);
      const bool in_l = __builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00008000u, 0x00040000u, 16>(zs));
      const double l = in_l ? c8 / d : 0.0;
      c8 = in_l ? l : c8;
      const double u10 = readlane_d(c10, 12);
      c10 = fma(-l, u10, c10);
    }
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000800u, 0x00000000u, 17>(zs))) lu[rb[0] + 0] = c0;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00040000u, 0x00000000u, 18>(zs))) lu[rb[0] + 1] = c1;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000801u, 0x00000008u, 19>(zs))) lu[rb[0] + 2] = c2;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x01040001u, 0x00000000u, 20>(zs))) lu[rb[0] + 3] = c3;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000049u, 0x00000000u, 21>(zs))) lu[rb[0] + 4] = c4;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000008u, 0x00000000u, 22>(zs))) lu[rb[0] + 5] = c5;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00001249u, 0x00000000u, 23>(zs))) lu[rb[0] + 6] = c6;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000200u, 0x00000000u, 24>(zs))) lu[rb[0] + 7] = c7;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00009240u, 0x00040000u, 25>(zs))) lu[rb[0] + 8] = c8;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00008000u, 0x00000000u, 26>(zs))) lu[rb[0] + 9] = c9;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00009000u, 0x00040000u, 27>(zs))) lu[rb[0] + 10] = c10;
  }

  // Solves: every lane loads its entry (in bounds for any row and column) a
  // chunk of terms ahead, and a select on the column's constant lane mask (an
  // SGPR pair) zeroes it outside the column's rows, off the substitution chain.
  // Forward substitution, column oriented, in level order: x_k = w_k is final
  // when its level is reached; rows of L(:, k) subtract L(i,k) x_k.
  static __device__ __forceinline__ void forward(const double* __restrict__ lu, const int (&rb)[1], double (&wv)[1], int lane) {
    (void)lane;
    double pf[5];  // factor entries, loaded 6 to 12 terms ahead of their update
    int zs = 0;  // laundered once per call: see lane_mask
    asm volatile("" : "+s"(zs));
    pf[0] = lu[rb[0] + 2];
    pf[1] = lu[rb[0] + 3];
    pf[2] = lu[rb[0] + 4];
    pf[3] = lu[rb[0] + 6];
    pf[4] = lu[rb[0] + 8];
    __builtin_amdgcn_sched_barrier(0);
    {  // level 0: pivots 2, 3
      const double x2 = readlane_d(wv[0], 35);
      const double x3 = readlane_d(wv[0], 24);
      { const double l = lane_keep(lane_mask<0x00000001u, 0x00000000u, 1>(zs), pf[0]);
        wv[0] = fma(-l, x2, wv[0]); }
      { const double l = lane_keep(lane_mask<0x00000001u, 0x00000000u, 2>(zs), pf[1]);
        wv[0] = fma(-l, x3, wv[0]); }
    }
    {  // level 1: pivots 4
      const double x4 = readlane_d(wv[0], 0);
      { const double l = lane_keep(lane_mask<0x00000048u, 0x00000000u, 3>(zs), pf[2]);
        wv[0] = fma(-l, x4, wv[0]); }
    }
    {  // level 2: pivots 6
      const double x6 = readlane_d(wv[0], 6);
      { const double l = lane_keep(lane_mask<0x00001200u, 0x00000000u, 4>(zs), pf[3]);
        wv[0] = fma(-l, x6, wv[0]); }
    }
    {  // level 3: pivots 8
      const double x8 = readlane_d(wv[0], 12);
      { const double l = lane_keep(lane_mask<0x00008000u, 0x00040000u, 5>(zs), pf[4]);
        wv[0] = fma(-l, x8, wv[0]); }
    }
  }

  // Backward substitution, column oriented, in level order: x_k = w_k / U(k,k)
  // (by the reciprocal diagonal), then rows of U(:, k) above it subtract U(i,k) x_k.
  static __device__ __forceinline__ void backward(const double* __restrict__ lu, const int (&rb)[1], double (&wv)[1], const double (&rdiag)[1], double (&x)[1], int lane) {
    // w_k is final once column k is reached (later columns update rows above
    // it only), so x = w * (1 / U(k,k)) is formed for every row at the end.
    (void)lane;
    double pf[5];  // factor entries, loaded 6 to 12 terms ahead of their update
    int zs = 0;  // laundered once per call: see lane_mask
    asm volatile("" : "+s"(zs));
    pf[0] = lu[rb[0] + 10];
    pf[1] = lu[rb[0] + 3];
    pf[2] = lu[rb[0] + 2];
    pf[3] = lu[rb[0] + 8];
    pf[4] = lu[rb[0] + 6];
    __builtin_amdgcn_sched_barrier(0);
    {  // level 0: pivots 10, 3, 2
      const double x10 = readlane_d(wv[0] * rdiag[0], 50);
      const double x3 = readlane_d(wv[0] * rdiag[0], 24);
      const double x2 = readlane_d(wv[0] * rdiag[0], 35);
      { const double u = lane_keep(lane_mask<0x00009000u, 0x00000000u, 1>(zs), pf[0]);
        wv[0] = fma(-u, x10, wv[0]); }
      { const double u = lane_keep(lane_mask<0x00040000u, 0x00000000u, 2>(zs), pf[1]);
        wv[0] = fma(-u, x3, wv[0]); }
      { const double u = lane_keep(lane_mask<0x00000800u, 0x00000000u, 3>(zs), pf[2]);
        wv[0] = fma(-u, x2, wv[0]); }
    }
    {  // level 1: pivots 8
      const double x8 = readlane_d(wv[0] * rdiag[0], 12);
      { const double u = lane_keep(lane_mask<0x00000240u, 0x00000000u, 4>(zs), pf[3]);
        wv[0] = fma(-u, x8, wv[0]); }
    }
    {  // level 2: pivots 6
      const double x6 = readlane_d(wv[0] * rdiag[0], 6);
      { const double u = lane_keep(lane_mask<0x00000009u, 0x00000000u, 5>(zs), pf[4]);
        wv[0] = fma(-u, x6, wv[0]); }
    }
    x[0] = wv[0] * rdiag[0];
  }

  // U^T z = b, column oriented over U's rows in level order: z_k = b_k / U(k,k)
  // once column k is reached, then every j > k in row k of U subtracts
  // U(k,j) z_k.  The lane holding row j (rv) reads U(k,j) at R[k] + j.
  static __device__ __forceinline__ void tforward(const double* __restrict__ lu, const int (&rv)[1], double (&b)[1], const double (&rdiag)[1], double (&z)[1], int lane) {
    // b_k is final once column k is reached, so z = b * (1 / U(k,k)) is formed
    // for every row at the end.
    {  // level 0: pivots 0, 1, 4, 5, 7, 9
      const double x0 = readlane_d(b[0] * rdiag[0], 11);
      const double x1 = readlane_d(b[0] * rdiag[0], 18);
      const double x4 = readlane_d(b[0] * rdiag[0], 0);
      const double x5 = readlane_d(b[0] * rdiag[0], 3);
      const double x7 = readlane_d(b[0] * rdiag[0], 9);
      const double x9 = readlane_d(b[0] * rdiag[0], 15);
      { const double u = lu[11 + rv[0]] * ((lane == 35) ? 1.0 : 0.0);
        b[0] = fma(-u, x0, b[0]); }
      { const double u = lu[18 + rv[0]] * ((lane == 24) ? 1.0 : 0.0);
        b[0] = fma(-u, x1, b[0]); }
      { const double u = lu[0 + rv[0]] * ((lane == 6) ? 1.0 : 0.0);
        b[0] = fma(-u, x4, b[0]); }
      { const double u = lu[3 + rv[0]] * ((lane == 6) ? 1.0 : 0.0);
        b[0] = fma(-u, x5, b[0]); }
      { const double u = lu[9 + rv[0]] * ((lane == 12) ? 1.0 : 0.0);
        b[0] = fma(-u, x7, b[0]); }
      { const double u = lu[15 + rv[0]] * ((lane == 50) ? 1.0 : 0.0);
        b[0] = fma(-u, x9, b[0]); }
    }
    {  // level 1: pivots 6
      const double x6 = readlane_d(b[0] * rdiag[0], 6);
      { const double u = lu[6 + rv[0]] * ((lane == 12) ? 1.0 : 0.0);
        b[0] = fma(-u, x6, b[0]); }
    }
    {  // level 2: pivots 8
      const double x8 = readlane_d(b[0] * rdiag[0], 12);
      { const double u = lu[12 + rv[0]] * ((lane == 50) ? 1.0 : 0.0);
        b[0] = fma(-u, x8, b[0]); }
    }
    z[0] = b[0] * rdiag[0];
  }

  // L^T x = z in place (unit diagonal), in descending level order: x_k = z_k
  // when column k is reached; every j < k in row k of L subtracts L(k,j) x_k.
  static __device__ __forceinline__ void tbackward(const double* __restrict__ lu, const int (&rv)[1], double (&z)[1], int lane) {
    {  // level 0: pivots 10, 9, 7, 5
      const double x10 = readlane_d(z[0], 50);
      const double x9 = readlane_d(z[0], 15);
      const double x7 = readlane_d(z[0], 9);
      const double x5 = readlane_d(z[0], 3);
      { const double l = lu[18 + rv[0]] * ((lane == 12) ? 1.0 : 0.0);
        z[0] = fma(-l, x10, z[0]); }
      { const double l = lu[15 + rv[0]] * ((lane == 12) ? 1.0 : 0.0);
        z[0] = fma(-l, x9, z[0]); }
      { const double l = lu[9 + rv[0]] * ((lane == 6) ? 1.0 : 0.0);
        z[0] = fma(-l, x7, z[0]); }
      { const double l = lu[3 + rv[0]] * ((lane == 0) ? 1.0 : 0.0);
        z[0] = fma(-l, x5, z[0]); }
    }
    {  // level 1: pivots 8
      const double x8 = readlane_d(z[0], 12);
      { const double l = lu[12 + rv[0]] * ((lane == 6) ? 1.0 : 0.0);
        z[0] = fma(-l, x8, z[0]); }
    }
    {  // level 2: pivots 6
      const double x6 = readlane_d(z[0], 6);
      { const double l = lu[6 + rv[0]] * ((lane == 0) ? 1.0 : 0.0);
        z[0] = fma(-l, x6, z[0]); }
    }
    {  // level 3: pivots 4
      const double x4 = readlane_d(z[0], 0);
      { const double l = lu[0 + rv[0]] * (((0x0000000801000000ull >> lane) & 1ull) ? 1.0 : 0.0);
        z[0] = fma(-l, x4, z[0]); }
    }
  }
};

} } }  // namespace mkpp::gen::carbon_wave
