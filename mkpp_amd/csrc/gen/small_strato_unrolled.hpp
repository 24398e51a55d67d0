// Generated by mkpp_amd.gen.emit_unrolled: dense unrolled LU and solves for small_strato
// (validation backend).  Do not edit: regenerate with `python -m mkpp_amd.gen`.
// The includer defines MKPP_UNROLLED_FN (function qualifiers).
#pragma once

namespace mkpp { namespace gen { namespace small_strato_unrolled {

struct Unrolled {
  static constexpr int N = 7;
  static constexpr int NLU = 32;
  // physical species at solver position i
  static MKPP_UNROLLED_FN int perm(int i) {
    constexpr int t[] = {5, 1, 0, 6, 2, 3, 4};
    return t[i];
  }

  // W = I / (gamma h) - J at the W locations, solver order, dense w[i * N + j];
  // J is the physical dense Jacobian J[r + N * c] (dF_r / dy_c)
  static MKPP_UNROLLED_FN void assemble_w(const double* J, double inv_g_dt, double* w) {
    w[0] = inv_g_dt;
    w[7] = -J[36];
    w[8] = inv_g_dt - J[8];
    w[11] = -J[15];
    w[14] = -J[35];
    w[15] = -J[7];
    w[16] = inv_g_dt - J[0];
    w[17] = -J[42];
    w[18] = -J[14];
    w[20] = -J[28];
    w[22] = -J[13];
    w[23] = -J[6];
    w[24] = inv_g_dt - J[48];
    w[25] = -J[20];
    w[26] = -J[27];
    w[27] = -J[34];
    w[29] = -J[9];
    w[30] = -J[2];
    w[31] = -J[44];
    w[32] = inv_g_dt - J[16];
    w[33] = -J[23];
    w[37] = -J[3];
    w[39] = -J[17];
    w[40] = inv_g_dt - J[24];
    w[41] = -J[31];
    w[44] = -J[4];
    w[46] = -J[18];
    w[47] = -J[25];
    w[48] = inv_g_dt - J[32];
  }

  static MKPP_UNROLLED_FN void factorize_chunk0(const double* w, double* lu) {
    lu[0] = w[0];
    lu[7] = w[7] / lu[0];
    lu[14] = w[14] / lu[0];
    lu[8] = w[8];
    lu[11] = w[11];
    lu[15] = w[15] / lu[8];
    lu[22] = w[22] / lu[8];
    lu[29] = w[29] / lu[8];
    lu[16] = w[16];
    lu[17] = w[17];
    lu[18] = (w[18] - lu[15] * lu[11]);
    lu[20] = w[20];
    lu[23] = w[23] / lu[16];
    lu[30] = w[30] / lu[16];
    lu[37] = w[37] / lu[16];
    lu[44] = w[44] / lu[16];
    lu[24] = (w[24] - lu[23] * lu[17]);
    lu[25] = ((w[25] - lu[22] * lu[11]) - lu[23] * lu[18]);
    lu[26] = w[26];
    lu[27] = (w[27] - lu[23] * lu[20]);
    lu[31] = (w[31] - lu[30] * lu[17]) / lu[24];
    lu[38] = (0.0 - lu[37] * lu[17]) / lu[24];
    lu[45] = (0.0 - lu[44] * lu[17]) / lu[24];
    lu[32] = (((w[32] - lu[29] * lu[11]) - lu[30] * lu[18]) - lu[31] * lu[25]);
    lu[33] = (w[33] - lu[31] * lu[26]);
    lu[34] = ((0.0 - lu[30] * lu[20]) - lu[31] * lu[27]);
    lu[39] = ((w[39] - lu[37] * lu[18]) - lu[38] * lu[25]) / lu[32];
    lu[46] = ((w[46] - lu[44] * lu[18]) - lu[45] * lu[25]) / lu[32];
    lu[40] = ((w[40] - lu[38] * lu[26]) - lu[39] * lu[33]);
    lu[41] = (((w[41] - lu[37] * lu[20]) - lu[38] * lu[27]) - lu[39] * lu[34]);
    lu[47] = ((w[47] - lu[45] * lu[26]) - lu[46] * lu[33]) / lu[40];
    lu[48] = ((((w[48] - lu[44] * lu[20]) - lu[45] * lu[27]) - lu[46] * lu[34]) - lu[47] * lu[41]);
  }
  // every LU entry in the symbolic order; lu[] must start zeroed
  static MKPP_UNROLLED_FN void factorize(const double* w, double* lu) {
    factorize_chunk0(w, lu);
  }

  // L y = rhs, then U x = y; solver order in and out
  static MKPP_UNROLLED_FN void solve(const double* lu, const double* rhs, double* x) {
    double y[7];
    y[0] = rhs[0];
    y[1] = (rhs[1] - lu[7] * y[0]);
    y[2] = ((rhs[2] - lu[14] * y[0]) - lu[15] * y[1]);
    y[3] = ((rhs[3] - lu[22] * y[1]) - lu[23] * y[2]);
    y[4] = (((rhs[4] - lu[29] * y[1]) - lu[30] * y[2]) - lu[31] * y[3]);
    y[5] = (((rhs[5] - lu[37] * y[2]) - lu[38] * y[3]) - lu[39] * y[4]);
    y[6] = ((((rhs[6] - lu[44] * y[2]) - lu[45] * y[3]) - lu[46] * y[4]) - lu[47] * y[5]);
    x[6] = y[6] / lu[48];
    x[5] = (y[5] - lu[41] * x[6]) / lu[40];
    x[4] = ((y[4] - lu[33] * x[5]) - lu[34] * x[6]) / lu[32];
    x[3] = (((y[3] - lu[25] * x[4]) - lu[26] * x[5]) - lu[27] * x[6]) / lu[24];
    x[2] = (((y[2] - lu[17] * x[3]) - lu[18] * x[4]) - lu[20] * x[6]) / lu[16];
    x[1] = (y[1] - lu[11] * x[4]) / lu[8];
    x[0] = y[0] / lu[0];
  }
};

} } }  // namespace mkpp::gen::small_strato_unrolled
