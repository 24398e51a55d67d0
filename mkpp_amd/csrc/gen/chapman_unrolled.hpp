// Generated by mkpp_amd.gen.emit_unrolled: dense unrolled LU and solves for chapman
// (validation backend).  Do not edit: regenerate with `python -m mkpp_amd.gen`.
// The includer defines MKPP_UNROLLED_FN (function qualifiers).
#pragma once

namespace mkpp { namespace gen { namespace chapman_unrolled {

struct Unrolled {
  static constexpr int N = 4;
  static constexpr int NLU = 13;
  // physical species at solver position i
  static MKPP_UNROLLED_FN int perm(int i) {
    constexpr int t[] = {3, 2, 0, 1};
    return t[i];
  }

  // W = I / (gamma h) - J at the W locations, solver order, dense w[i * N + j];
  // J is the physical dense Jacobian J[r + N * c] (dF_r / dy_c)
  static MKPP_UNROLLED_FN void assemble_w(const double* J, double inv_g_dt, double* w) {
    w[0] = inv_g_dt;
    w[4] = -J[14];
    w[5] = inv_g_dt - J[10];
    w[6] = -J[2];
    w[7] = -J[6];
    w[8] = -J[12];
    w[9] = -J[8];
    w[10] = inv_g_dt - J[0];
    w[11] = -J[4];
    w[12] = -J[13];
    w[13] = -J[9];
    w[14] = -J[1];
    w[15] = inv_g_dt - J[5];
  }

  static MKPP_UNROLLED_FN void factorize_chunk0(const double* w, double* lu) {
    lu[0] = w[0];
    lu[4] = w[4] / lu[0];
    lu[8] = w[8] / lu[0];
    lu[12] = w[12] / lu[0];
    lu[5] = w[5];
    lu[6] = w[6];
    lu[7] = w[7];
    lu[9] = w[9] / lu[5];
    lu[13] = w[13] / lu[5];
    lu[10] = (w[10] - lu[9] * lu[6]);
    lu[11] = (w[11] - lu[9] * lu[7]);
    lu[14] = (w[14] - lu[13] * lu[6]) / lu[10];
    lu[15] = ((w[15] - lu[13] * lu[7]) - lu[14] * lu[11]);
  }
  // every LU entry in the symbolic order; lu[] must start zeroed
  static MKPP_UNROLLED_FN void factorize(const double* w, double* lu) {
    factorize_chunk0(w, lu);
  }

  // L y = rhs, then U x = y; solver order in and out
  static MKPP_UNROLLED_FN void solve(const double* lu, const double* rhs, double* x) {
    double y[4];
    y[0] = rhs[0];
    y[1] = (rhs[1] - lu[4] * y[0]);
    y[2] = ((rhs[2] - lu[8] * y[0]) - lu[9] * y[1]);
    y[3] = (((rhs[3] - lu[12] * y[0]) - lu[13] * y[1]) - lu[14] * y[2]);
    x[3] = y[3] / lu[15];
    x[2] = (y[2] - lu[11] * x[3]) / lu[10];
    x[1] = ((y[1] - lu[6] * x[2]) - lu[7] * x[3]) / lu[5];
    x[0] = y[0] / lu[0];
  }
};

} } }  // namespace mkpp::gen::chapman_unrolled
