// Generated by mkpp_amd.gen.emit_hip from mkpp_amd/mechanisms/carbon/mechanism.json.
// Do not edit: regenerate with `python -m mkpp_amd.gen`.
#pragma once
#include "../mkpp_device.hpp"

namespace mkpp { namespace gen { namespace carbon {

struct Mech {
  static constexpr int N = 11;        // species, physical order
  static constexpr int NV = 11;       // variable species (logical RCM order)
  static constexpr int NF = 1;       // fixed species (>= 1 for array sizing)
  static constexpr int NFIXED = 0;
  static constexpr int NPHOTO = 1;
  static constexpr int NUM_RUNTIME_PARAMETERS = 0;
  static constexpr int NG = 5;
  // rate-constant slots: the flux groups, then Kp(T) of each equilibrium block
  static constexpr int NK = 5;
  static constexpr bool USES_CLOUD_WATER = false;
  static constexpr int NLU = 27;
  static constexpr int NJ = 22;
  static constexpr int NUM_REACTIONS = 5;
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = true;
  static constexpr int THIRD_BODY = -1;
  static constexpr const char* NAME = "carbon";
  static constexpr const char* SPECIES[N] = {"CH4", "CO", "PCOfromCH4", "PCOfromNMVOC", "LCH4byOH", "LCH4byCl", "LCObyOH", "FixedOH", "FixedCl", "DummyCH4", "DummyNMVOC"};

  // logical variable position -> physical species
  static __host__ __device__ constexpr int var_phys(int v) {
    constexpr int t[] = {3, 2, 10, 9, 1, 6, 7, 4, 0, 5, 8};
    return t[v];
  }
  static __host__ __device__ constexpr int fixed_phys(int q) {
    constexpr int t[] = {-1};
    return t[q];
  }
  // reduction variant: W locations (LU slot, row, column); tolerance index of a row
  static constexpr int NWM = 27;
  static __host__ __device__ constexpr int wm_slot(int e) {
    constexpr int t[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 18, 19, 20, 21, 22, 23, 24, 25, 26};
    return t[e];
  }
  static __host__ __device__ constexpr int wm_row(int e) {
    constexpr int t[] = {0, 0, 1, 1, 2, 3, 4, 4, 4, 4, 5, 5, 5, 6, 6, 6, 7, 7, 7, 8, 8, 8, 9, 9, 9, 10, 10};
    return t[e];
  }
  static __host__ __device__ constexpr int wm_col(int e) {
    constexpr int t[] = {0, 2, 1, 3, 2, 3, 2, 3, 4, 6, 4, 5, 6, 4, 6, 8, 6, 7, 8, 6, 8, 10, 8, 9, 10, 8, 10};
    return t[e];
  }
  static __host__ __device__ constexpr int tol_index(int v) {
    constexpr int t[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10};
    return t[v];
  }

  // k_r and dk_r/d[M] per flux group (src/mkpp/lowering.py:199-313); clw: the
  // cell's cloud liquid water (cloud-gated heterogeneous reactions)
  static __device__ __forceinline__ void rate_constants(double T, double M, const double* __restrict__ jv,
                                                        double clw, double* __restrict__ k,
                                                        double* __restrict__ dk) {
    (void)T; (void)M; (void)jv; (void)clw;
    k[0] = (((2.45e-12 * exp(((-1775.0) / T)))) * 1.0);
    dk[0] = 0.0;
    k[1] = (((9.6e-12 * exp(((-1360.0) / T)))) * 1.0);
    dk[1] = 0.0;
    k[2] = (4.2566446e-15 * 1.0);
    dk[2] = 0.0;
    k[3] = (7.3679649e-14 * 1.0);
    dk[3] = 0.0;
    k[4] = (38199.012 * 1.0);
    dk[4] = 0.0;
  }

  // F over variable species in logical order; x: variable state (logical), yf: fixed species
  static __device__ __forceinline__ void rates(const double* __restrict__ x, const double* __restrict__ yf,
                                               const double* __restrict__ k, double* __restrict__ f) {
    (void)yf;
    const double fl0 = (k[0] * (1.0 * x[8] * 1.0));
    const double fl1 = (k[1] * (1.0 * x[8] * 1.0));
    const double fl2 = (k[2] * (1.0 * 1.0));
    const double fl3 = (k[3] * (1.0 * x[4] * 1.0));
    const double fl4 = (k[4] * (1.0 * 1.0));
    f[0] = (0.0 + (1.0 * fl4));
    f[1] = (0.0 + (1.0 * fl2));
    f[2] = (0.0 + (-1.0 * fl4));
    f[3] = (0.0 + (-1.0 * fl2));
    f[4] = (((0.0 + (1.0 * fl2)) + (-1.0 * fl3)) + (1.0 * fl4));
    f[5] = (0.0 + (1.0 * fl3));
    f[6] = ((0.0 + (-1.0 * fl0)) + (-1.0 * fl3));
    f[7] = (0.0 + (1.0 * fl0));
    f[8] = ((0.0 + (-1.0 * fl0)) + (-1.0 * fl1));
    f[9] = (0.0 + (1.0 * fl1));
    f[10] = (0.0 + (-1.0 * fl1));
  }

  // W = I/(gamma h) - J for variable rows/columns, written into LU storage (integrate.j2:103-107)
  static __device__ __forceinline__ void jac_w(const double* __restrict__ x, const double* __restrict__ yf,
                                               const double* __restrict__ k, const double* __restrict__ dk,
                                               double inv_g_dt, double* __restrict__ lu) {
    (void)yf; (void)dk;
    const double d0_0 = (0.0 + (k[0] * 1.0 * 1.0 * 1.0));
    const double d0_7 = (0.0 + (k[0] * 1.0 * 1.0 * x[8]));
    const double d1_0 = (0.0 + (k[1] * 1.0 * 1.0 * 1.0));
    const double d1_8 = (0.0 + (k[1] * 1.0 * 1.0 * x[8]));
    const double d2_9 = (0.0 + (k[2] * 1.0 * 1.0));
    const double d3_1 = (0.0 + (k[3] * 1.0 * 1.0 * 1.0));
    const double d3_7 = (0.0 + (k[3] * 1.0 * 1.0 * x[4]));
    const double d4_10 = (0.0 + (k[4] * 1.0 * 1.0));
    lu[0] = inv_g_dt;
    lu[1] = 0.0 - (0.0 + (1.0 * d4_10));
    lu[2] = inv_g_dt;
    lu[3] = 0.0 - (0.0 + (1.0 * d2_9));
    lu[4] = inv_g_dt - (0.0 + (-1.0 * d4_10));
    lu[5] = inv_g_dt - (0.0 + (-1.0 * d2_9));
    lu[6] = 0.0 - (0.0 + (1.0 * d4_10));
    lu[7] = 0.0 - (0.0 + (1.0 * d2_9));
    lu[8] = inv_g_dt - (0.0 + (-1.0 * d3_1));
    lu[9] = 0.0 - (0.0 + (-1.0 * d3_7));
    lu[10] = 0.0 - (0.0 + (1.0 * d3_1));
    lu[11] = inv_g_dt;
    lu[12] = 0.0 - (0.0 + (1.0 * d3_7));
    lu[13] = 0.0 - (0.0 + (-1.0 * d3_1));
    lu[14] = inv_g_dt - ((0.0 + (-1.0 * d0_7)) + (-1.0 * d3_7));
    lu[15] = 0.0 - (0.0 + (-1.0 * d0_0));
    lu[16] = 0.0 - (0.0 + (1.0 * d0_7));
    lu[17] = inv_g_dt;
    lu[18] = 0.0 - (0.0 + (1.0 * d0_0));
    lu[19] = 0.0 - (0.0 + (-1.0 * d0_7));
    lu[20] = inv_g_dt - ((0.0 + (-1.0 * d0_0)) + (-1.0 * d1_0));
    lu[21] = 0.0 - (0.0 + (-1.0 * d1_8));
    lu[22] = 0.0 - (0.0 + (1.0 * d1_0));
    lu[23] = inv_g_dt;
    lu[24] = 0.0 - (0.0 + (1.0 * d1_8));
    lu[25] = 0.0 - (0.0 + (-1.0 * d1_0));
    lu[26] = inv_g_dt - (0.0 + (-1.0 * d1_8));
  }

  // compact Doolittle factorization in the reference operation order (compact_plan.py:6-70)
  static __device__ __forceinline__ void factor(double* __restrict__ lu) {
    lu[6] = lu[6] / lu[4];
    lu[7] = lu[7] / lu[5];
    lu[10] = lu[10] / lu[8];
    lu[13] = lu[13] / lu[8];
    lu[12] = (lu[12] - lu[10] * lu[9]);
    lu[14] = (lu[14] - lu[13] * lu[9]);
    lu[16] = lu[16] / lu[14];
    lu[19] = lu[19] / lu[14];
    lu[18] = (lu[18] - lu[16] * lu[15]);
    lu[20] = (lu[20] - lu[19] * lu[15]);
    lu[22] = lu[22] / lu[20];
    lu[25] = lu[25] / lu[20];
    lu[24] = (lu[24] - lu[22] * lu[21]);
    lu[26] = (lu[26] - lu[25] * lu[21]);
  }

  // forward/backward substitution (supernodal_solve.hpp.j2:10-42), logical variable order
  static __device__ __forceinline__ void solve(const double* __restrict__ lu, const double* __restrict__ rhs,
                                               double* __restrict__ x) {
    double wk[11];
    wk[0] = rhs[0];
    wk[1] = rhs[1];
    wk[2] = rhs[2];
    wk[3] = rhs[3];
    wk[4] = ((rhs[4] - lu[6] * wk[2]) - lu[7] * wk[3]);
    wk[5] = (rhs[5] - lu[10] * wk[4]);
    wk[6] = (rhs[6] - lu[13] * wk[4]);
    wk[7] = (rhs[7] - lu[16] * wk[6]);
    wk[8] = (rhs[8] - lu[19] * wk[6]);
    wk[9] = (rhs[9] - lu[22] * wk[8]);
    wk[10] = (rhs[10] - lu[25] * wk[8]);
    x[10] = wk[10] / lu[26];
    x[9] = (wk[9] - lu[24] * x[10]) / lu[23];
    x[8] = (wk[8] - lu[21] * x[10]) / lu[20];
    x[7] = (wk[7] - lu[18] * x[8]) / lu[17];
    x[6] = (wk[6] - lu[15] * x[8]) / lu[14];
    x[5] = (wk[5] - lu[12] * x[6]) / lu[11];
    x[4] = (wk[4] - lu[9] * x[6]) / lu[8];
    x[3] = wk[3] / lu[5];
    x[2] = wk[2] / lu[4];
    x[1] = (wk[1] - lu[3] * x[3]) / lu[2];
    x[0] = (wk[0] - lu[1] * x[2]) / lu[0];
  }

  // physical-order F (fixed species rows are zero)
  static __device__ __forceinline__ void rates_full(const double* __restrict__ y, const double* __restrict__ k,
                                                    double* __restrict__ f) {
    const double fl0 = (k[0] * (1.0 * y[0] * 1.0));
    const double fl1 = (k[1] * (1.0 * y[0] * 1.0));
    const double fl2 = (k[2] * (1.0 * 1.0));
    const double fl3 = (k[3] * (1.0 * y[1] * 1.0));
    const double fl4 = (k[4] * (1.0 * 1.0));
    f[0] = ((0.0 + (-1.0 * fl0)) + (-1.0 * fl1));
    f[1] = (((0.0 + (1.0 * fl2)) + (-1.0 * fl3)) + (1.0 * fl4));
    f[2] = (0.0 + (1.0 * fl2));
    f[3] = (0.0 + (1.0 * fl4));
    f[4] = (0.0 + (1.0 * fl0));
    f[5] = (0.0 + (1.0 * fl1));
    f[6] = (0.0 + (1.0 * fl3));
    f[7] = ((0.0 + (-1.0 * fl0)) + (-1.0 * fl3));
    f[8] = (0.0 + (-1.0 * fl1));
    f[9] = (0.0 + (-1.0 * fl2));
    f[10] = (0.0 + (-1.0 * fl4));
  }

  // dense J[i + N*j] including fixed-species columns; structural zeros are written as 0.0
  static __device__ __forceinline__ void jac_full(const double* __restrict__ y, const double* __restrict__ k,
                                                  const double* __restrict__ dk, double* __restrict__ J,
                                                  long stride) {
    (void)dk;
    for (int e = 0; e < 121; ++e) J[e * stride] = 0.0;
    J[0L * stride] = ((0.0 + (-1.0 * (0.0 + (k[0] * 1.0 * 1.0 * 1.0)))) + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * 1.0))));
    J[77L * stride] = (0.0 + (-1.0 * (0.0 + (k[0] * 1.0 * 1.0 * y[0]))));
    J[88L * stride] = (0.0 + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0]))));
    J[12L * stride] = (0.0 + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * 1.0))));
    J[78L * stride] = (0.0 + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[1]))));
    J[100L * stride] = (0.0 + (1.0 * (0.0 + (k[2] * 1.0 * 1.0))));
    J[111L * stride] = (0.0 + (1.0 * (0.0 + (k[4] * 1.0 * 1.0))));
    J[101L * stride] = (0.0 + (1.0 * (0.0 + (k[2] * 1.0 * 1.0))));
    J[113L * stride] = (0.0 + (1.0 * (0.0 + (k[4] * 1.0 * 1.0))));
    J[4L * stride] = (0.0 + (1.0 * (0.0 + (k[0] * 1.0 * 1.0 * 1.0))));
    J[81L * stride] = (0.0 + (1.0 * (0.0 + (k[0] * 1.0 * 1.0 * y[0]))));
    J[5L * stride] = (0.0 + (1.0 * (0.0 + (k[1] * 1.0 * 1.0 * 1.0))));
    J[93L * stride] = (0.0 + (1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0]))));
    J[17L * stride] = (0.0 + (1.0 * (0.0 + (k[3] * 1.0 * 1.0 * 1.0))));
    J[83L * stride] = (0.0 + (1.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[1]))));
    J[7L * stride] = (0.0 + (-1.0 * (0.0 + (k[0] * 1.0 * 1.0 * 1.0))));
    J[18L * stride] = (0.0 + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * 1.0))));
    J[84L * stride] = ((0.0 + (-1.0 * (0.0 + (k[0] * 1.0 * 1.0 * y[0])))) + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[1]))));
    J[8L * stride] = (0.0 + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * 1.0))));
    J[96L * stride] = (0.0 + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0]))));
    J[108L * stride] = (0.0 + (-1.0 * (0.0 + (k[2] * 1.0 * 1.0))));
    J[120L * stride] = (0.0 + (-1.0 * (0.0 + (k[4] * 1.0 * 1.0))));
  }
};

} } }  // namespace mkpp::gen::carbon
