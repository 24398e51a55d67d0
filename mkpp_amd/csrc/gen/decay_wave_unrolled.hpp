// Generated by mkpp_amd.gen.emit_unrolled: dense unrolled LU and solves for decay_wave
// (validation backend).  Do not edit: regenerate with `python -m mkpp_amd.gen`.
// The includer defines MKPP_UNROLLED_FN (function qualifiers).
#pragma once

namespace mkpp { namespace gen { namespace decay_wave_unrolled {

struct Unrolled {
  static constexpr int N = 1;
  static constexpr int NLU = 1;
  // physical species at solver position i
  static MKPP_UNROLLED_FN int perm(int i) {
    constexpr int t[] = {0};
    return t[i];
  }

  // W = I / (gamma h) - J at the W locations, solver order, dense w[i * N + j];
  // J is the physical dense Jacobian J[r + N * c] (dF_r / dy_c)
  static MKPP_UNROLLED_FN void assemble_w(const double* J, double inv_g_dt, double* w) {
    w[0] = inv_g_dt - J[0];
  }

  static MKPP_UNROLLED_FN void factorize_chunk0(const double* w, double* lu) {
    lu[0] = w[0];
  }
  // every LU entry in the symbolic order; lu[] must start zeroed
  static MKPP_UNROLLED_FN void factorize(const double* w, double* lu) {
    factorize_chunk0(w, lu);
  }

  // L y = rhs, then U x = y; solver order in and out
  static MKPP_UNROLLED_FN void solve(const double* lu, const double* rhs, double* x) {
    double y[1];
    y[0] = rhs[0];
    x[0] = y[0] / lu[0];
  }
};

} } }  // namespace mkpp::gen::decay_wave_unrolled
