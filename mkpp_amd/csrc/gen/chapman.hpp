// Generated by mkpp_amd.gen.emit_hip from mkpp_amd/mechanisms/chapman/mechanism.json.
// Do not edit: regenerate with `python -m mkpp_amd.gen`.
#pragma once
#include "../mkpp_device.hpp"

namespace mkpp { namespace gen { namespace chapman {

struct Mech {
  static constexpr int N = 4;        // species, physical order
  static constexpr int NV = 3;       // variable species (logical RCM order)
  static constexpr int NF = 1;       // fixed species (>= 1 for array sizing)
  static constexpr int NFIXED = 1;
  static constexpr int NPHOTO = 2;
  static constexpr int NUM_RUNTIME_PARAMETERS = 2;
  static constexpr int NG = 4;
  // rate-constant slots: the flux groups, then Kp(T) of each equilibrium block
  static constexpr int NK = 4;
  static constexpr bool USES_CLOUD_WATER = false;
  static constexpr int NLU = 9;
  static constexpr int NJ = 12;
  static constexpr int NUM_REACTIONS = 4;
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = false;
  static constexpr int THIRD_BODY = 3;
  static constexpr const char* NAME = "chapman";
  static constexpr const char* SPECIES[N] = {"O", "O2", "O3", "M"};

  // logical variable position -> physical species
  static __host__ __device__ constexpr int var_phys(int v) {
    constexpr int t[] = {2, 0, 1};
    return t[v];
  }
  static __host__ __device__ constexpr int fixed_phys(int q) {
    constexpr int t[] = {3};
    return t[q];
  }
  // reduction variant: W locations (LU slot, row, column); tolerance index of a row
  static constexpr int NWM = 9;
  static __host__ __device__ constexpr int wm_slot(int e) {
    constexpr int t[] = {0, 1, 2, 3, 4, 5, 6, 7, 8};
    return t[e];
  }
  static __host__ __device__ constexpr int wm_row(int e) {
    constexpr int t[] = {0, 0, 0, 1, 1, 1, 2, 2, 2};
    return t[e];
  }
  static __host__ __device__ constexpr int wm_col(int e) {
    constexpr int t[] = {0, 1, 2, 0, 1, 2, 0, 1, 2};
    return t[e];
  }
  static __host__ __device__ constexpr int tol_index(int v) {
    constexpr int t[] = {1, 2, 3};
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
    k[1] = (6e-34 * 1.0);
    dk[1] = 0.0;
    k[2] = jv[1];
    dk[2] = 0.0;
    k[3] = (8e-12 * 1.0);
    dk[3] = 0.0;
  }

  // F over variable species in logical order; x: variable state (logical), yf: fixed species
  static __device__ __forceinline__ void rates(const double* __restrict__ x, const double* __restrict__ yf,
                                               const double* __restrict__ k, double* __restrict__ f) {
    (void)yf;
    const double fl0 = (k[0] * (1.0 * x[2]));
    const double fl1 = (k[1] * (1.0 * x[1] * x[2] * yf[0]));
    const double fl2 = (k[2] * (1.0 * x[0]));
    const double fl3 = (k[3] * (1.0 * x[1] * x[0]));
    f[0] = (((0.0 + (1.0 * fl1)) + (-1.0 * fl2)) + (-1.0 * fl3));
    f[1] = ((((0.0 + (2.0 * fl0)) + (-1.0 * fl1)) + (1.0 * fl2)) + (-1.0 * fl3));
    f[2] = ((((0.0 + (-1.0 * fl0)) + (-1.0 * fl1)) + (1.0 * fl2)) + (2.0 * fl3));
  }

  // W = I/(gamma h) - J for variable rows/columns, written into LU storage (integrate.j2:103-107)
  static __device__ __forceinline__ void jac_w(const double* __restrict__ x, const double* __restrict__ yf,
                                               const double* __restrict__ k, const double* __restrict__ dk,
                                               double inv_g_dt, double* __restrict__ lu) {
    (void)yf; (void)dk;
    const double d0_1 = (0.0 + (k[0] * 1.0 * 1.0));
    const double d1_0 = (0.0 + (k[1] * 1.0 * 1.0 * x[2] * yf[0]));
    const double d1_1 = (0.0 + (k[1] * 1.0 * 1.0 * x[1] * yf[0]));
    const double d2_2 = (0.0 + (k[2] * 1.0 * 1.0));
    const double d3_0 = (0.0 + (k[3] * 1.0 * 1.0 * x[0]));
    const double d3_2 = (0.0 + (k[3] * 1.0 * 1.0 * x[1]));
    lu[0] = inv_g_dt - ((0.0 + (-1.0 * d2_2)) + (-1.0 * d3_2));
    lu[1] = 0.0 - ((0.0 + (1.0 * d1_0)) + (-1.0 * d3_0));
    lu[2] = 0.0 - (0.0 + (1.0 * d1_1));
    lu[3] = 0.0 - ((0.0 + (1.0 * d2_2)) + (-1.0 * d3_2));
    lu[4] = inv_g_dt - ((0.0 + (-1.0 * d1_0)) + (-1.0 * d3_0));
    lu[5] = 0.0 - ((0.0 + (2.0 * d0_1)) + (-1.0 * d1_1));
    lu[6] = 0.0 - ((0.0 + (1.0 * d2_2)) + (2.0 * d3_2));
    lu[7] = 0.0 - ((0.0 + (-1.0 * d1_0)) + (2.0 * d3_0));
    lu[8] = inv_g_dt - ((0.0 + (-1.0 * d0_1)) + (-1.0 * d1_1));
  }

  // compact Doolittle factorization in the reference operation order (compact_plan.py:6-70)
  static __device__ __forceinline__ void factor(double* __restrict__ lu) {
    lu[3] = lu[3] / lu[0];
    lu[6] = lu[6] / lu[0];
    lu[4] = (lu[4] - lu[3] * lu[1]);
    lu[5] = (lu[5] - lu[3] * lu[2]);
    lu[7] = (lu[7] - lu[6] * lu[1]) / lu[4];
    lu[8] = ((lu[8] - lu[6] * lu[2]) - lu[7] * lu[5]);
  }

  // forward/backward substitution (supernodal_solve.hpp.j2:10-42), logical variable order
  static __device__ __forceinline__ void solve(const double* __restrict__ lu, const double* __restrict__ rhs,
                                               double* __restrict__ x) {
    double wk[3];
    wk[0] = rhs[0];
    wk[1] = (rhs[1] - lu[3] * wk[0]);
    wk[2] = ((rhs[2] - lu[6] * wk[0]) - lu[7] * wk[1]);
    x[2] = wk[2] / lu[8];
    x[1] = (wk[1] - lu[5] * x[2]) / lu[4];
    x[0] = ((wk[0] - lu[1] * x[1]) - lu[2] * x[2]) / lu[0];
  }

  // physical-order F (fixed species rows are zero)
  static __device__ __forceinline__ void rates_full(const double* __restrict__ y, const double* __restrict__ k,
                                                    double* __restrict__ f) {
    const double fl0 = (k[0] * (1.0 * y[1]));
    const double fl1 = (k[1] * (1.0 * y[0] * y[1] * y[3]));
    const double fl2 = (k[2] * (1.0 * y[2]));
    const double fl3 = (k[3] * (1.0 * y[0] * y[2]));
    f[0] = ((((0.0 + (2.0 * fl0)) + (-1.0 * fl1)) + (1.0 * fl2)) + (-1.0 * fl3));
    f[1] = ((((0.0 + (-1.0 * fl0)) + (-1.0 * fl1)) + (1.0 * fl2)) + (2.0 * fl3));
    f[2] = (((0.0 + (1.0 * fl1)) + (-1.0 * fl2)) + (-1.0 * fl3));
    f[3] = 0.0;
  }

  // dense J[i + N*j] including fixed-species columns; structural zeros are written as 0.0
  static __device__ __forceinline__ void jac_full(const double* __restrict__ y, const double* __restrict__ k,
                                                  const double* __restrict__ dk, double* __restrict__ J,
                                                  long stride) {
    (void)dk;
    for (int e = 0; e < 16; ++e) J[e * stride] = 0.0;
    J[0L * stride] = ((0.0 + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[1] * y[3])))) + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[2]))));
    J[4L * stride] = ((0.0 + (2.0 * (0.0 + (k[0] * 1.0 * 1.0)))) + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0] * y[3]))));
    J[8L * stride] = ((0.0 + (1.0 * (0.0 + (k[2] * 1.0 * 1.0)))) + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[0]))));
    J[12L * stride] = (0.0 + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0] * y[1]))));
    J[1L * stride] = ((0.0 + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[1] * y[3])))) + (2.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[2]))));
    J[5L * stride] = ((0.0 + (-1.0 * (0.0 + (k[0] * 1.0 * 1.0)))) + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0] * y[3]))));
    J[9L * stride] = ((0.0 + (1.0 * (0.0 + (k[2] * 1.0 * 1.0)))) + (2.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[0]))));
    J[13L * stride] = (0.0 + (-1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0] * y[1]))));
    J[2L * stride] = ((0.0 + (1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[1] * y[3])))) + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[2]))));
    J[6L * stride] = (0.0 + (1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0] * y[3]))));
    J[10L * stride] = ((0.0 + (-1.0 * (0.0 + (k[2] * 1.0 * 1.0)))) + (-1.0 * (0.0 + (k[3] * 1.0 * 1.0 * y[0]))));
    J[14L * stride] = (0.0 + (1.0 * (0.0 + (k[1] * 1.0 * 1.0 * y[0] * y[1]))));
  }
};

} } }  // namespace mkpp::gen::chapman
