// Generated by mkpp_amd.gen.emit_wave: straight-line LU and solves for chapman.
// Do not edit: regenerate with `python -m mkpp_amd.gen`.
#pragma once
#include "chapman_wave.hpp"

namespace mkpp { namespace gen { namespace chapman_wave {

__device__ const uint32_t __attribute__((aligned(16))) CF_TERM[256] = {16908801, 50594051, 50594564, 50594564, 16842752, 16974338, 50594564, 50594564, 16843008, 197122, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564, 50594564};
__device__ const uint16_t __attribute__((aligned(8))) CF_PART[256] = {5, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 6, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 7, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4, 4};
__device__ const uint32_t CW_TERM[128] = {65540, 131073, 131076, 65537, 0, 131076, 65537, 65536, 131074, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 65542, 65541, 65542, 65541, 65538, 6, 5, 65538, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615, 196615};
__device__ const uint16_t CW_ADDR[64] = {0, 1, 3, 4, 5, 6, 7, 8, 2, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9, 9};
__device__ const uint32_t __attribute__((aligned(16))) CS_TERM[256] = {393220, 458759, 458759, 458759, 327681, 458759, 458759, 458759, 131072, 458759, 458759, 458759, 458755, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759, 458759};
__device__ const uint16_t __attribute__((aligned(8))) CS_PART[256] = {0, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 1, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 2, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64, 64};
__device__ const uint16_t CS_FIX[1] = {3};
__device__ const uint16_t CG_TERM[256] = {1, 257, 513, 257, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 258, 258, 514, 2, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 512, 256, 256, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771, 771};
__device__ const uint16_t CG_GRP[64] = {0, 1, 2, 3, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535, 65535};

struct Code {
  static constexpr int TF = 8, NFP = 1, TW = 2, NWP = 1;
  static constexpr int FBITS = 8;  // flux-slot bits of a 16-bit F term (coefficient above)
  // F partial slots scr[PBASE, PBASE + chunks); a row reads FMAXC[bank] of its FMAXCP-wide list
  static constexpr int PBASE = 5, FSLOTS = 8, FMAXCP = 4;
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
    // Rows and columns 0..2 in registers: lane = row (bank 0), c<j> =
    // the row's entry in column j (0 where the pattern has none).  Pivot rows
    // are broadcast with readlane; the multiplier is zeroed outside L(k), so
    // every update is an unconditional fma that leaves other lanes unchanged.
    double c0 = lane_keep(lane_mask<0x00000049u, 0x00000000u, 1>(zs), lu[rb[0] + 0]);
    double c1 = lane_keep(lane_mask<0x00000049u, 0x00000000u, 2>(zs), lu[rb[0] + 1]);
    double c2 = lane_keep(lane_mask<0x00000049u, 0x00000000u, 3>(zs), lu[rb[0] + 2]);
    {  // pivot 0: 2 rows x 2 columns
      const double d = readlane_d(c0, 0);
      const bool in_l = __builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000048u, 0x00000000u, 4>(zs));
      const double l = in_l ? c0 / d : 0.0;
      c0 = in_l ? l : c0;
      const double u1 = readlane_d(c1, 0);
      const double u2 = readlane_d(c2, 0);
      c1 = fma(-l, u1, c1);
      c2 = fma(-l, u2, c2);
    }
    {  // pivot 1: 1 rows x 1 columns
      const double d = readlane_d(c1, 3);
      const bool in_l = __builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000040u, 0x00000000u, 5>(zs));
      const double l = in_l ? c1 / d : 0.0;
      c1 = in_l ? l : c1;
      const double u2 = readlane_d(c2, 3);
      c2 = fma(-l, u2, c2);
    }
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000049u, 0x00000000u, 6>(zs))) lu[rb[0] + 0] = c0;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000049u, 0x00000000u, 7>(zs))) lu[rb[0] + 1] = c1;
    if (__builtin_amdgcn_inverse_ballot_w64(lane_mask<0x00000049u, 0x00000000u, 8>(zs))) lu[rb[0] + 2] = c2;
  }

  // Solves: every lane loads its entry (in bounds for any row and column) a
  // chunk of terms ahead, and a select on the column's constant lane mask (an
  // SGPR pair) zeroes it outside the column's rows, off the substitution chain.
  // Forward substitution, column oriented, in level order: x_k = w_k is final
  // when its level is reached; rows of L(:, k) subtract L(i,k) x_k.
  static __device__ __forceinline__ void forward(const double* __restrict__ lu, const int (&rb)[1], double (&wv)[1], int lane) {
    (void)lane;
    double pf[2];  // factor entries, loaded 6 to 12 terms ahead of their update
    int zs = 0;  // laundered once per call: see lane_mask
    asm volatile("" : "+s"(zs));
    pf[0] = lu[rb[0] + 0];
    pf[1] = lu[rb[0] + 1];
    __builtin_amdgcn_sched_barrier(0);
    {  // level 0: pivots 0
      const double x0 = readlane_d(wv[0], 0);
      { const double l = lane_keep(lane_mask<0x00000048u, 0x00000000u, 1>(zs), pf[0]);
        wv[0] = fma(-l, x0, wv[0]); }
    }
    {  // level 1: pivots 1
      const double x1 = readlane_d(wv[0], 3);
      { const double l = lane_keep(lane_mask<0x00000040u, 0x00000000u, 2>(zs), pf[1]);
        wv[0] = fma(-l, x1, wv[0]); }
    }
  }

  // Backward substitution, column oriented, in level order: x_k = w_k / U(k,k)
  // (by the reciprocal diagonal), then rows of U(:, k) above it subtract U(i,k) x_k.
  static __device__ __forceinline__ void backward(const double* __restrict__ lu, const int (&rb)[1], double (&wv)[1], const double (&rdiag)[1], double (&x)[1], int lane) {
    // w_k is final once column k is reached (later columns update rows above
    // it only), so x = w * (1 / U(k,k)) is formed for every row at the end.
    (void)lane;
    double pf[2];  // factor entries, loaded 6 to 12 terms ahead of their update
    int zs = 0;  // laundered once per call: see lane_mask
    asm volatile("" : "+s"(zs));
    pf[0] = lu[rb[0] + 2];
    pf[1] = lu[rb[0] + 1];
    __builtin_amdgcn_sched_barrier(0);
    {  // level 0: pivots 2
      const double x2 = readlane_d(wv[0] * rdiag[0], 6);
      { const double u = lane_keep(lane_mask<0x00000009u, 0x00000000u, 1>(zs), pf[0]);
        wv[0] = fma(-u, x2, wv[0]); }
    }
    {  // level 1: pivots 1
      const double x1 = readlane_d(wv[0] * rdiag[0], 3);
      { const double u = lane_keep(lane_mask<0x00000001u, 0x00000000u, 2>(zs), pf[1]);
        wv[0] = fma(-u, x1, wv[0]); }
    }
    x[0] = wv[0] * rdiag[0];
  }

  // U^T z = b, column oriented over U's rows in level order: z_k = b_k / U(k,k)
  // once column k is reached, then every j > k in row k of U subtracts
  // U(k,j) z_k.  The lane holding row j (rv) reads U(k,j) at R[k] + j.
  static __device__ __forceinline__ void tforward(const double* __restrict__ lu, const int (&rv)[1], double (&b)[1], const double (&rdiag)[1], double (&z)[1], int lane) {
    // b_k is final once column k is reached, so z = b * (1 / U(k,k)) is formed
    // for every row at the end.
    {  // level 0: pivots 0
      const double x0 = readlane_d(b[0] * rdiag[0], 0);
      { const double u = lu[0 + rv[0]] * (((0x0000000000000048ull >> lane) & 1ull) ? 1.0 : 0.0);
        b[0] = fma(-u, x0, b[0]); }
    }
    {  // level 1: pivots 1
      const double x1 = readlane_d(b[0] * rdiag[0], 3);
      { const double u = lu[3 + rv[0]] * ((lane == 6) ? 1.0 : 0.0);
        b[0] = fma(-u, x1, b[0]); }
    }
    z[0] = b[0] * rdiag[0];
  }

  // L^T x = z in place (unit diagonal), in descending level order: x_k = z_k
  // when column k is reached; every j < k in row k of L subtracts L(k,j) x_k.
  static __device__ __forceinline__ void tbackward(const double* __restrict__ lu, const int (&rv)[1], double (&z)[1], int lane) {
    {  // level 0: pivots 2
      const double x2 = readlane_d(z[0], 6);
      { const double l = lu[6 + rv[0]] * (((0x0000000000000009ull >> lane) & 1ull) ? 1.0 : 0.0);
        z[0] = fma(-l, x2, z[0]); }
    }
    {  // level 1: pivots 1
      const double x1 = readlane_d(z[0], 3);
      { const double l = lu[3 + rv[0]] * ((lane == 0) ? 1.0 : 0.0);
        z[0] = fma(-l, x1, z[0]); }
    }
  }
};

} } }  // namespace mkpp::gen::chapman_wave
