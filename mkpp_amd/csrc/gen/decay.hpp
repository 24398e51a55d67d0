// Generated by mkpp_amd.gen.emit_hip from mkpp_amd/mechanisms/decay/mechanism.json.
// Do not edit: regenerate with `python -m mkpp_amd.gen`.
#pragma once
#include "../mkpp_device.hpp"

namespace mkpp { namespace gen { namespace decay {

struct Mech {
  static constexpr int N = 1;        // species, physical order
  static constexpr int NV = 1;       // variable species (logical RCM order)
  static constexpr int NF = 1;       // fixed species (>= 1 for array sizing)
  static constexpr int NFIXED = 0;
  static constexpr int NPHOTO = 1;
  static constexpr int NUM_RUNTIME_PARAMETERS = 1;
  static constexpr int NG = 1;
  // rate-constant slots: the flux groups, then Kp(T) of each equilibrium block
  static constexpr int NK = 1;
  static constexpr bool USES_CLOUD_WATER = false;
  static constexpr int NLU = 1;
  static constexpr int NJ = 1;
  static constexpr int NUM_REACTIONS = 1;
  static constexpr bool SUPPORTS_RUNTIME_ENVIRONMENT = false;
  static constexpr int THIRD_BODY = -1;
  static constexpr const char* NAME = "decay";
  static constexpr const char* SPECIES[N] = {"X"};

  // logical variable position -> physical species
  static __host__ __device__ constexpr int var_phys(int v) {
    constexpr int t[] = {0};
    return t[v];
  }
  static __host__ __device__ constexpr int fixed_phys(int q) {
    constexpr int t[] = {-1};
    return t[q];
  }
  // reduction variant: W locations (LU slot, row, column); tolerance index of a row
  static constexpr int NWM = 1;
  static __host__ __device__ constexpr int wm_slot(int e) {
    constexpr int t[] = {0};
    return t[e];
  }
  static __host__ __device__ constexpr int wm_row(int e) {
    constexpr int t[] = {0};
    return t[e];
  }
  static __host__ __device__ constexpr int wm_col(int e) {
    constexpr int t[] = {0};
    return t[e];
  }
  static __host__ __device__ constexpr int tol_index(int v) {
    constexpr int t[] = {0};
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
  }

  // F over variable species in logical order; x: variable state (logical), yf: fixed species
  static __device__ __forceinline__ void rates(const double* __restrict__ x, const double* __restrict__ yf,
                                               const double* __restrict__ k, double* __restrict__ f) {
    (void)yf;
    const double fl0 = (k[0] * (1.0 * x[0]));
    f[0] = (0.0 + (-1.0 * fl0));
  }

  // W = I/(gamma h) - J for variable rows/columns, written into LU storage (integrate.j2:103-107)
  static __device__ __forceinline__ void jac_w(const double* __restrict__ x, const double* __restrict__ yf,
                                               const double* __restrict__ k, const double* __restrict__ dk,
                                               double inv_g_dt, double* __restrict__ lu) {
    (void)yf; (void)dk;
    const double d0_0 = (0.0 + (k[0] * 1.0 * 1.0));
    lu[0] = inv_g_dt - (0.0 + (-1.0 * d0_0));
  }

  // compact Doolittle factorization in the reference operation order (compact_plan.py:6-70)
  static __device__ __forceinline__ void factor(double* __restrict__ lu) {
  }

  // forward/backward substitution (supernodal_solve.hpp.j2:10-42), logical variable order
  static __device__ __forceinline__ void solve(const double* __restrict__ lu, const double* __restrict__ rhs,
                                               double* __restrict__ x) {
    double wk[1];
    wk[0] = rhs[0];
    x[0] = wk[0] / lu[0];
  }

  // physical-order F (fixed species rows are zero)
  static __device__ __forceinline__ void rates_full(const double* __restrict__ y, const double* __restrict__ k,
                                                    double* __restrict__ f) {
    const double fl0 = (k[0] * (1.0 * y[0]));
    f[0] = (0.0 + (-1.0 * fl0));
  }

  // dense J[i + N*j] including fixed-species columns; structural zeros are written as 0.0
  static __device__ __forceinline__ void jac_full(const double* __restrict__ y, const double* __restrict__ k,
                                                  const double* __restrict__ dk, double* __restrict__ J,
                                                  long stride) {
    (void)dk;
    for (int e = 0; e < 1; ++e) J[e * stride] = 0.0;
    J[0L * stride] = (0.0 + (-1.0 * (0.0 + (k[0] * 1.0 * 1.0))));
  }
};

} } }  // namespace mkpp::gen::decay
