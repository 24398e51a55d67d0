// Generated by mkpp_amd.gen.emit_unrolled: dense unrolled LU and solves for carbon
// (validation backend).  Do not edit: regenerate with `python -m mkpp_amd.gen`.
// The includer defines MKPP_UNROLLED_FN (function qualifiers).
#pragma once

namespace mkpp { namespace gen { namespace carbon_unrolled {

struct Unrolled {
  static constexpr int N = 11;
  static constexpr int NLU = 27;
  // physical species at solver position i
  static MKPP_UNROLLED_FN int perm(int i) {
    constexpr int t[] = {3, 2, 10, 9, 1, 6, 7, 4, 0, 5, 8};
    return t[i];
  }

  // W = I / (gamma h) - J at the W locations, solver order, dense w[i * N + j];
  // J is the physical dense Jacobian J[r + N * c] (dF_r / dy_c)
  static MKPP_UNROLLED_FN void assemble_w(const double* J, double inv_g_dt, double* w) {
    w[0] = inv_g_dt;
    w[2] = -J[113];
    w[12] = inv_g_dt;
    w[14] = -J[101];
    w[24] = inv_g_dt - J[120];
    w[36] = inv_g_dt - J[108];
    w[46] = -J[111];
    w[47] = -J[100];
    w[48] = inv_g_dt - J[12];
    w[50] = -J[78];
    w[59] = -J[17];
    w[60] = inv_g_dt;
    w[61] = -J[83];
    w[70] = -J[18];
    w[72] = inv_g_dt - J[84];
    w[74] = -J[7];
    w[83] = -J[81];
    w[84] = inv_g_dt;
    w[85] = -J[4];
    w[94] = -J[77];
    w[96] = inv_g_dt - J[0];
    w[98] = -J[88];
    w[107] = -J[5];
    w[108] = inv_g_dt;
    w[109] = -J[93];
    w[118] = -J[8];
    w[120] = inv_g_dt - J[96];
  }

  static MKPP_UNROLLED_FN void factorize_chunk0(const double* w, double* lu) {
    lu[0] = w[0];
    lu[2] = w[2];
    lu[12] = w[12];
    lu[14] = w[14];
    lu[24] = w[24];
    lu[46] = w[46] / lu[24];
    lu[36] = w[36];
    lu[47] = w[47] / lu[36];
    lu[48] = w[48];
    lu[50] = w[50];
    lu[59] = w[59] / lu[48];
    lu[70] = w[70] / lu[48];
    lu[60] = w[60];
    lu[61] = (w[61] - lu[59] * lu[50]);
    lu[72] = (w[72] - lu[70] * lu[50]);
    lu[74] = w[74];
    lu[83] = w[83] / lu[72];
    lu[94] = w[94] / lu[72];
    lu[84] = w[84];
    lu[85] = (w[85] - lu[83] * lu[74]);
    lu[96] = (w[96] - lu[94] * lu[74]);
    lu[98] = w[98];
    lu[107] = w[107] / lu[96];
    lu[118] = w[118] / lu[96];
    lu[108] = w[108];
    lu[109] = (w[109] - lu[107] * lu[98]);
    lu[120] = (w[120] - lu[118] * lu[98]);
  }
  // every LU entry in the symbolic order; lu[] must start zeroed
  static MKPP_UNROLLED_FN void factorize(const double* w, double* lu) {
    factorize_chunk0(w, lu);
  }

  // L y = rhs, then U x = y; solver order in and out
  static MKPP_UNROLLED_FN void solve(const double* lu, const double* rhs, double* x) {
    double y[11];
    y[0] = rhs[0];
    y[1] = rhs[1];
    y[2] = rhs[2];
    y[3] = rhs[3];
    y[4] = ((rhs[4] - lu[46] * y[2]) - lu[47] * y[3]);
    y[5] = (rhs[5] - lu[59] * y[4]);
    y[6] = (rhs[6] - lu[70] * y[4]);
    y[7] = (rhs[7] - lu[83] * y[6]);
    y[8] = (rhs[8] - lu[94] * y[6]);
    y[9] = (rhs[9] - lu[107] * y[8]);
    y[10] = (rhs[10] - lu[118] * y[8]);
    x[10] = y[10] / lu[120];
    x[9] = (y[9] - lu[109] * x[10]) / lu[108];
    x[8] = (y[8] - lu[98] * x[10]) / lu[96];
    x[7] = (y[7] - lu[85] * x[8]) / lu[84];
    x[6] = (y[6] - lu[74] * x[8]) / lu[72];
    x[5] = (y[5] - lu[61] * x[6]) / lu[60];
    x[4] = (y[4] - lu[50] * x[6]) / lu[48];
    x[3] = y[3] / lu[36];
    x[2] = y[2] / lu[24];
    x[1] = (y[1] - lu[14] * x[3]) / lu[12];
    x[0] = (y[0] - lu[2] * x[2]) / lu[0];
  }
};

} } }  // namespace mkpp::gen::carbon_unrolled
