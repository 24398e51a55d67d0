// Generated by mkpp_amd.gen.emit_hip from mkpp_amd/mechanisms/small_strato/mechanism.json.
// Do not edit: regenerate with `python -m mkpp_amd.gen`.
#pragma once
#include "../mkpp_device.hpp"

namespace mkpp { namespace gen { namespace small_strato {

struct Mech {
  static constexpr int N = 7;        // species, physical order
  static constexpr int NV = 6;       // variable species (logical RCM order)
  static constexpr int NF = 1;       // fixed species (>= 1 for array sizing)
  static constexpr int NFIXED = 1;
  static constexpr int NPHOTO = 4;
  static constexpr int NUM_RUNTIME_PARAMETERS = 4;
  static constexpr int NG = 10;
  // rate-constant slots: the flux groups, then Kp(T) of each equilibrium block
  static constexpr int NK = 10;
  static constexpr bool USES_CLOUD_WATER = false;
  static constexpr int NLU = 29;
  static constexpr int NJ = 28;
  static constexpr int NUM_REACTIONS = 10;
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = false;
  static constexpr int THIRD_BODY = 5;
  static constexpr const char* NAME = "small_strato";
  static constexpr const char* SPECIES[N] = {"O", "O1D", "O3", "NO", "NO2", "M", "O2"};

  // logical variable position -> physical species
  static __host__ __device__ constexpr int var_phys(int v) {
    constexpr int t[] = {1, 0, 6, 2, 3, 4};
    return t[v];
  }
  static __host__ __device__ constexpr int fixed_phys(int q) {
    constexpr int t[] = {5};
    return t[q];
  }
  // reduction variant: W locations (LU slot, row, column); tolerance index of a row
  static constexpr int NWM = 26;
  static __host__ __device__ constexpr int wm_slot(int e) {
    constexpr int t[] = {0, 1, 2, 3, 4, 5, 6, 7, 8, 9, 10, 11, 12, 13, 14, 15, 16, 17, 19, 21, 22, 23, 24, 26, 27, 28};
    return t[e];
  }
  static __host__ __device__ constexpr int wm_row(int e) {
    constexpr int t[] = {0, 0, 1, 1, 1, 1, 1, 2, 2, 2, 2, 2, 2, 3, 3, 3, 3, 3, 4, 4, 4, 4, 5, 5, 5, 5};
    return t[e];
  }
  static __host__ __device__ constexpr int wm_col(int e) {
    constexpr int t[] = {0, 3, 0, 1, 2, 3, 5, 0, 1, 2, 3, 4, 5, 0, 1, 2, 3, 4, 1, 3, 4, 5, 1, 3, 4, 5};
    return t[e];
  }
  static __host__ __device__ constexpr int tol_index(int v) {
    constexpr int t[] = {1, 2, 3, 4, 5, 6};
    return t[v];
  }

  // k_r and dk_r/d[M] per flux group (src/mkpp/lowering.py:199-313); clw: the
  // cell's cloud liquid water (cloud-gated heterogeneous reactions)
  static __device__ __forceinline__ void rate_constants(double T, double M, const double* __restrict__ jv,
                                                        double clw, double* __restrict__ k,
                                                        double* __restrict__ dk) {
    (void)T; (void)M; (void)jv; (void)clw;
    k[0] = jv[0];
    dk[0] = 0.0;
    k[1] = (8.018e-17 * 1.0);
    dk[1] = 0.0;
    k[2] = jv[1];
    dk[2] = 0.0;
    k[3] = (1.576e-15 * 1.0);
    dk[3] = 0.0;
    k[4] = jv[2];
    dk[4] = 0.0;
    k[5] = (7.11e-11 * 1.0);
    dk[5] = 0.0;
    k[6] = (1.2e-10 * 1.0);
    dk[6] = 0.0;
    k[7] = (6.062e-15 * 1.0);
    dk[7] = 0.0;
    k[8] = (1.069e-11 * 1.0);
    dk[8] = 0.0;
    k[9] = jv[3];
    dk[9] = 0.0;
  }

  // F over variable species in logical order; x: variable state (logical), yf: fixed species
  static __device__ __forceinline__ void rates(const double* __restrict__ x, const double* __restrict__ yf,
                                               const double* __restrict__ k, double* __restrict__ f) {
    (void)yf;
    const double fl0 = (k[0] * (1.0 * x[2]));
    const double fl1 = (k[1] * (1.0 * x[1] * x[2]));
    const double fl2 = (k[2] * (1.0 * x[3]));
    const double fl3 = (k[3] * (1.0 * x[1] * x[3]));
    const double fl4 = (k[4] * (1.0 * x[3]));
    const double fl5 = (k[5] * (1.0 * x[0] * yf[0]));
    const double fl6 = (k[6] * (1.0 * x[0] * x[3]));
    const double fl7 = (k[7] * (1.0 * x[3] * x[4]));
    const double fl8 = (k[8] * (1.0 * x[1] * x[5]));
    const double fl9 = (k[9] * (1.0 * x[5]));
    f[0] = (((0.0 + (1.0 * fl4)) + (-1.0 * fl5)) + (-1.0 * fl6));
    f[1] = (((((((0.0 + (2.0 * fl0)) + (-1.0 * fl1)) + (1.0 * fl2)) + (-1.0 * fl3)) + (1.0 * fl5)) + (-1.0 * fl8)) + (1.0 * fl9));
    f[2] = ((((((((0.0 + (-1.0 * fl0)) + (-1.0 * fl1)) + (1.0 * fl2)) + (2.0 * fl3)) + (1.0 * fl4)) + (2.0 * fl6)) + (1.0 * fl7)) + (1.0 * fl8));
    f[3] = ((((((0.0 + (1.0 * fl1)) + (-1.0 * fl2)) + (-1.0 * fl3)) + (-1.0 * fl4)) + (-1.0 * fl6)) + (-1.0 * fl7));
    f[4] = (((0.0 + (-1.0 * fl7)) + (1.0 * fl8)) + (1.0 * fl9));
    f[5] = (((0.0 + (1.0 * fl7)) + (-1.0 * fl8)) + (-1.0 * fl9));
  }

  // W = I/(gamma h) - J for variable rows/columns, written into LU storage (integrate.j2:103-107)
  static __device__ __forceinline__ void jac_w(const double* __restrict__ x, const double* __restrict__ yf,
                                               const double* __restrict__ k, const double* __restrict__ dk,
                                               double inv_g_dt, double* __restrict__ lu) {
    (void)yf; (void)dk;
    const double d0_6 = (0.0 + (k[0] * 1.0 * 1.0));
    const double d1_0 = (0.0 + (k[1] * 1.0 * 1.0 * x[2]));
    const double d1_6 = (0.0 + (k[1] * 1.0 * 1.0 * x[1]));
    const double d2_2 = (0.0 + (k[2] * 1.0 * 1.0));
    const double d3_0 = (0.0 + (k[3] * 1.0 * 1.0 * x[3]));
    const double d3_2 = (0.0 + (k[3] * 1.0 * 1.0 * x[1]));
    const double d4_2 = (0.0 + (k[4] * 1.0 * 1.0));
    const double d5_1 = (0.0 + (k[5] * 1.0 * 1.0 * yf[0]));
    const double d6_1 = (0.0 + (k[6] * 1.0 * 1.0 * x[3]));
    const double d6_2 = (0.0 + (k[6] * 1.0 * 1.0 * x[0]));
    const double d7_2 = (0.0 + (k[7] * 1.0 * 1.0 * x[4]));
    const double d7_3 = (0.0 + (k[7] * 1.0 * 1.0 * x[3]));
    const double d8_0 = (0.0 + (k[8] * 1.0 * 1.0 * x[5]));
    const double d8_4 = (0.0 + (k[8] * 1.0 * 1.0 * x[1]));
    const double d9_4 = (0.0 + (k[9] * 1.0 * 1.0));
    lu[0] = inv_g_dt - ((0.0 + (-1.0 * d5_1)) + (-1.0 * d6_1));
    lu[1] = 0.0 - ((0.0 + (1.0 * d4_2)) + (-1.0 * d6_2));
    lu[2] = 0.0 - (0.0 + (1.0 * d5_1));
    lu[3] = inv_g_dt - (((0.0 + (-1.0 * d1_0)) + (-1.0 * d3_0)) + (-1.0 * d8_0));
    lu[4] = 0.0 - ((0.0 + (2.0 * d0_6)) + (-1.0 * d1_6));
    lu[5] = 0.0 - ((0.0 + (1.0 * d2_2)) + (-1.0 * d3_2));
    lu[6] = 0.0 - ((0.0 + (-1.0 * d8_4)) + (1.0 * d9_4));
    lu[7] = 0.0 - (0.0 + (2.0 * d6_1));
    lu[8] = 0.0 - (((0.0 + (-1.0 * d1_0)) + (2.0 * d3_0)) + (1.0 * d8_0));
    lu[9] = inv_g_dt - ((0.0 + (-1.0 * d0_6)) + (-1.0 * d1_6));
    lu[10] = 0.0 - (((((0.0 + (1.0 * d2_2)) + (2.0 * d3_2)) + (1.0 * d4_2)) + (2.0 * d6_2)) + (1.0 * d7_2));
    lu[11] = 0.0 - (0.0 + (1.0 * d7_3));
    lu[12] = 0.0 - (0.0 + (1.0 * d8_4));
    lu[13] = 0.0 - (0.0 + (-1.0 * d6_1));
    lu[14] = 0.0 - ((0.0 + (1.0 * d1_0)) + (-1.0 * d3_0));
    lu[15] = 0.0 - (0.0 + (1.0 * d1_6));
    lu[16] = inv_g_dt - (((((0.0 + (-1.0 * d2_2)) + (-1.0 * d3_2)) + (-1.0 * d4_2)) + (-1.0 * d6_2)) + (-1.0 * d7_2));
    lu[17] = 0.0 - (0.0 + (-1.0 * d7_3));
    lu[18] = 0.0;
    lu[19] = 0.0 - (0.0 + (1.0 * d8_0));
    lu[20] = 0.0;
    lu[21] = 0.0 - (0.0 + (-1.0 * d7_2));
    lu[22] = inv_g_dt - (0.0 + (-1.0 * d7_3));
    lu[23] = 0.0 - ((0.0 + (1.0 * d8_4)) + (1.0 * d9_4));
    lu[24] = 0.0 - (0.0 + (-1.0 * d8_0));
    lu[25] = 0.0;
    lu[26] = 0.0 - (0.0 + (1.0 * d7_2));
    lu[27] = 0.0 - (0.0 + (1.0 * d7_3));
    lu[28] = inv_g_dt - ((0.0 + (-1.0 * d8_4)) + (-1.0 * d9_4));
  }

  // compact Doolittle factorization in the reference operation order (compact_plan.py:6-70)
  static __device__ __forceinline__ void factor(double* __restrict__ lu) {
    lu[2] = lu[2] / lu[0];
    lu[7] = lu[7] / lu[0];
    lu[13] = lu[13] / lu[0];
    lu[5] = (lu[5] - lu[2] * lu[1]);
    lu[8] = lu[8] / lu[3];
    lu[14] = lu[14] / lu[3];
    lu[19] = lu[19] / lu[3];
    lu[24] = lu[24] / lu[3];
    lu[9] = (lu[9] - lu[8] * lu[4]);
    lu[10] = ((lu[10] - lu[7] * lu[1]) - lu[8] * lu[5]);
    lu[12] = (lu[12] - lu[8] * lu[6]);
    lu[15] = (lu[15] - lu[14] * lu[4]) / lu[9];
    lu[20] = (lu[20] - lu[19] * lu[4]) / lu[9];
    lu[25] = (lu[25] - lu[24] * lu[4]) / lu[9];
    lu[16] = (((lu[16] - lu[13] * lu[1]) - lu[14] * lu[5]) - lu[15] * lu[10]);
    lu[17] = (lu[17] - lu[15] * lu[11]);
    lu[18] = ((lu[18] - lu[14] * lu[6]) - lu[15] * lu[12]);
    lu[21] = ((lu[21] - lu[19] * lu[5]) - lu[20] * lu[10]) / lu[16];
    lu[26] = ((lu[26] - lu[24] * lu[5]) - lu[25] * lu[10]) / lu[16];
    lu[22] = ((lu[22] - lu[20] * lu[11]) - lu[21] * lu[17]);
    lu[23] = (((lu[23] - lu[19] * lu[6]) - lu[20] * lu[12]) - lu[21] * lu[18]);
    lu[27] = ((lu[27] - lu[25] * lu[11]) - lu[26] * lu[17]) / lu[22];
    lu[28] = ((((lu[28] - lu[24] * lu[6]) - lu[25] * lu[12]) - lu[26] * lu[18]) - lu[27] * lu[23]);
  }

  // forward/backward substitution (supernodal_solve.hpp.j2:10-42), logical variable order
  static __device__ __forceinline__ void solve(const double* __restrict__ lu, const double* __restrict__ rhs,
                                               double* __restrict__ x) {
    double wk[6];
    wk[0] = rhs[0];
    wk[1] = (rhs[1] - lu[2] * wk[0]);
    wk[2] = ((rhs[2] - lu[7] * wk[0]) - lu[8] * wk[1]);
    wk[3] = (((rhs[3] - lu[13] * wk[0]) - lu[14] * wk[1]) - lu[15] * wk[2]);
    wk[4] = (((rhs[4] - lu[19] * wk[1]) - lu[20] * wk[2]) - lu[21] * wk[3]);
    wk[5] = ((((rhs[5] - lu[24] * wk[1]) - lu[25] * wk[2]) - lu[26] * wk[3]) - lu[27] * wk[4]);
    x[5] = wk[5] / lu[28];
    x[4] = (wk[4] - lu[23] * x[5]) / lu[22];
    x[3] = ((wk[3] - lu[17] * x[4]) - lu[18] * x[5]) / lu[16];
    x[2] = (((wk[2] - lu[10] * x[3]) - lu[11] * x[4]) - lu[12] * x[5]) / lu[9];
    x[1] = (((wk[1] - lu[4] * x[2]) - lu[5] * x[3]) - lu[6] * x[5]) / lu[3];
    x[0] = (wk[0] - lu[1] * x[3]) / lu[0];
  }

  // physical-order F (fixed species rows are zero)
  static __device__ __forceinline__ void rates_full(const double* __restrict__ y, const double* __restrict__ k,
                                                    double* __restrict__ f) {
    const double fl0 = (k[0] * (1.0 * y[6]));
    const double fl1 = (k[1] * (1.0 * y[0] * y[6]));
    const double fl2 = (k[2] * (1.0 * y[2]));
    const double fl3 = (k[3] * (1.0 * y[0] * y[2]));
    const double fl4 = (k[4] * (1.0 * y[2]));
    const double fl5 = (k[5] * (1.0 * y[1] * y[5]));
    const double fl6 = (k[6] * (1.0 * y[1] * y[2]));
    const double fl7 = (k[7] * (1.0 * y[2] * y[3]));
    const double fl8 = (k[8] * (1.0 * y[0] * y[4]));
    const double fl9 = (k[9] * (1.0 * y[4]));
    f[0] = (((((((0.0 + (2.0 * fl0)) + (-1.0 * fl1)) + (1.0 * fl2)) + (-1.0 * fl3)) + (1.0 * fl5)) + (-1.0 * fl8)) + (1.0 * fl9));
    f[1] = (((0.0 + (1.0 * fl4)) + (-1.0 * fl5)) + (-1.0 * fl6));
    f[2] = ((((((0.0 + (1.0 * fl1)) + (-1.0 * fl2)) + (-1.0 * fl3)) + (-1.0 * fl4)) + (-1.0 * fl6)) + (-1.0 * fl7));
    f[3] = (((0.0 + (-1.0 * fl7)) + (1.0 * fl8)) + (1.0 * fl9));
    f[4] = (((0.0 + (1.0 * fl7)) + (-1.0 * fl8)) + (-1.0 * fl9));
    f[5] = 0.0;
    f[6] = ((((((((0.0 + (-1.0 * fl0)) + (-1.0 * fl1)) + (1.0 * fl2)) + (2.0 * fl3)) + (1.0 * fl4)) + (2.0 * fl6)) + (1.0 * fl7)) + (1.0 * fl8));
  }

  // dense J[i + N*j] including fixed-species columns; structural zeros are written as 0.0
  static __device__ __forceinline__ void jac_full(const double* __restrict__ y, const double* __restrict__ k,
                                                  const double* __restrict__ dk, double* __restrict__ J,
                                                  long stride) {
    (void)dk;
    for (int e = 0; e < 49; ++e) J[e * stride] = 0.0;
    J[0L * stride] = (((0.0 + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[6])))) + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[2])))) + (-1.0 * (0.0 + (k[8] * 1.0 * 1.0 * y[4]))));
    J[7L * stride] = (0.0 + (1.0 * (0.0 + (k[5] * 1.0 * 1.0 * y[5]))));
    J[14L * stride] = ((0.0 + (1.0 * (0.0 + (k[2] * 1.0 * 1.0)))) + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[0]))));
    J[28L * stride] = ((0.0 + (-1.0 * (0.0 + (k[8] * 1.0 * 1.0 * y[0])))) + (1.0 * (0.0 + (k[9] * 1.0 * 1.0))));
    J[35L * stride] = (0.0 + (1.0 * (0.0 + (k[5] * 1.0 * 1.0 * y[1]))));
    J[42L * stride] = ((0.0 + (2.0 * (0.0 + (k[0] * 1.0 * 1.0)))) + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0]))));
    J[8L * stride] = ((0.0 + (-1.0 * (0.0 + (k[5] * 1.0 * 1.0 * y[5])))) + (-1.0 * (0.0 + (k[6] * 1.0 * 1.0 * y[2]))));
    J[15L * stride] = ((0.0 + (1.0 * (0.0 + (k[4] * 1.0 * 1.0)))) + (-1.0 * (0.0 + (k[6] * 1.0 * 1.0 * y[1]))));
    J[36L * stride] = (0.0 + (-1.0 * (0.0 + (k[5] * 1.0 * 1.0 * y[1]))));
    J[2L * stride] = ((0.0 + (1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[6])))) + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[2]))));
    J[9L * stride] = (0.0 + (-1.0 * (0.0 + (k[6] * 1.0 * 1.0 * y[2]))));
    J[16L * stride] = (((((0.0 + (-1.0 * (0.0 + (k[2] * 1.0 * 1.0)))) + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[0])))) + (-1.0 * (0.0 + (k[4] * 1.0 * 1.0)))) + (-1.0 * (0.0 + (k[6] * 1.0 * 1.0 * y[1])))) + (-1.0 * (0.0 + (k[7] * 1.0 * 1.0 * y[3]))));
    J[23L * stride] = (0.0 + (-1.0 * (0.0 + (k[7] * 1.0 * 1.0 * y[2]))));
    J[44L * stride] = (0.0 + (1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0]))));
    J[3L * stride] = (0.0 + (1.0 * (0.0 + (k[8] * 1.0 * 1.0 * y[4]))));
    J[17L * stride] = (0.0 + (-1.0 * (0.0 + (k[7] * 1.0 * 1.0 * y[3]))));
    J[24L * stride] = (0.0 + (-1.0 * (0.0 + (k[7] * 1.0 * 1.0 * y[2]))));
    J[31L * stride] = ((0.0 + (1.0 * (0.0 + (k[8] * 1.0 * 1.0 * y[0])))) + (1.0 * (0.0 + (k[9] * 1.0 * 1.0))));
    J[4L * stride] = (0.0 + (-1.0 * (0.0 + (k[8] * 1.0 * 1.0 * y[4]))));
    J[18L * stride] = (0.0 + (1.0 * (0.0 + (k[7] * 1.0 * 1.0 * y[3]))));
    J[25L * stride] = (0.0 + (1.0 * (0.0 + (k[7] * 1.0 * 1.0 * y[2]))));
    J[32L * stride] = ((0.0 + (-1.0 * (0.0 + (k[8] * 1.0 * 1.0 * y[0])))) + (-1.0 * (0.0 + (k[9] * 1.0 * 1.0))));
    J[6L * stride] = (((0.0 + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[6])))) + (2.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[2])))) + (1.0 * (0.0 + (k[8] * 1.0 * 1.0 * y[4]))));
    J[13L * stride] = (0.0 + (2.0 * (0.0 + (k[6] * 1.0 * 1.0 * y[2]))));
    J[20L * stride] = (((((0.0 + (1.0 * (0.0 + (k[2] * 1.0 * 1.0)))) + (2.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[0])))) + (1.0 * (0.0 + (k[4] * 1.0 * 1.0)))) + (2.0 * (0.0 + (k[6] * 1.0 * 1.0 * y[1])))) + (1.0 * (0.0 + (k[7] * 1.0 * 1.0 * y[3]))));
    J[27L * stride] = (0.0 + (1.0 * (0.0 + (k[7] * 1.0 * 1.0 * y[2]))));
    J[34L * stride] = (0.0 + (1.0 * (0.0 + (k[8] * 1.0 * 1.0 * y[0]))));
    J[48L * stride] = ((0.0 + (-1.0 * (0.0 + (k[0] * 1.0 * 1.0)))) + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0]))));
  }
};

} } }  // namespace mkpp::gen::small_strato
