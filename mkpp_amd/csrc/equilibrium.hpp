// Gas-aerosol partition of an EQUILIBRIUM block for the lane engine.
//
// An EQUILIBRIUM block adds a relaxation tau_inv (eq_i - C_i) to the tendency
// of each species the partition model returns (src/mkpp/lowering.py:374-414).
// The NH4_NO3_SO4 model (src/mkpp/equilibrium/nh4_no3_so4.py:94-280) maps the
// element totals CN = NH3 + NH4a, COx = HNO3 + NO3an1..3 and CS = SO2 + SO4
// to equilibrium NH3, NH4a, HNO3, NO3an1, NO3an2, NO3an3:
//
//   R = CN / (CS + 1e-30),  w_i = 1/2 (1 + u_i / sqrt(1 + u_i^2)),
//   u_1 = (R - 1) / width,  u_2 = (R - 2) / width,
//   e = CN - 2 CS,  X = 1/2 ((e + COx) - sqrt((e - COx)^2 + 4 Kp)),
//   sp(x) = 1/2 (x + sqrt(x^2 + 1e-20)),  X+ = sp(X),
//   NH3 = w_2 sp(e - X+),  NH4a = CN - NH3,  A = w_1 w_2 X+,
//   NO3an_k = A / 3,  HNO3 = COx - A,
//
// with Kp(T) = 4.39e-17 exp(-dH / Rgas (1/T - 1/298.15)), dH = -74735 J/mol,
// Rgas = 8.314 J/(mol K).  With DERIV the derivatives of the six values in
// (CN, COx, CS) are formed alongside, by the chain rule through the same
// intermediates, for the Jacobian, and dd[k] = d[k][own element] - 1 (NH3 and
// NH4a belong to CN, the others to COx), the diagonal of the relaxation's
// Jacobian over tau_inv, without the cancellation of 1 - 1.  Terms that
// cancel are taken in equivalent closed forms: for x < 0, sp(x) =
// 1e-20 / (2 (s - x)) and sp'(x) = 1e-20 / (2 s (s - x)), s = sqrt(x^2 + 1e-20);
// 1 -+ g / sqrt(g^2 + 4 Kp) = 4 Kp / (sqrt(.) (sqrt(.) +- g)) on the side where
// g = e - COx has the cancelling sign; X = 2 (e COx - Kp) / ((e + COx) +
// sqrt(.)) where e + COx > 0; e - X+ through e - X = (g + sqrt(.)) / 2.
#pragma once

#include <hip/hip_runtime.h>

namespace mkpp {

// Kp of NH3(g) + HNO3(g) <-> NH4NO3 at temperature T (van't Hoff form).
__device__ __forceinline__ double nh4no3_kp(double T) {
  return 4.39e-17 * exp((74735.0 / 8.314) * (1.0 / T - 1.0 / 298.15));
}

template <bool DERIV>
__device__ __forceinline__ void nh4_no3_so4_partition(double CN, double COx, double CS, double Kp, double inv_width,
                                                      double (&v)[6], double (&d)[6][3], double (&dd)[6]) {
  const double den = CS + 1e-30;
  const double R = CN / den;
  const double u1 = (R - 1.0) * inv_width, u2 = (R - 2.0) * inv_width;
  const double q1 = sqrt(1.0 + u1 * u1), q2 = sqrt(1.0 + u2 * u2);
  // w = (1 + u / q) / 2 and its complement 1 - w = (1 - u / q) / 2, each in the
  // form that does not cancel: 1 / (2 q (q -+ u))
  const double w1 = u1 >= 0.0 ? 0.5 * (1.0 + u1 / q1) : 0.5 / (q1 * (q1 - u1));
  const double w2 = u2 >= 0.0 ? 0.5 * (1.0 + u2 / q2) : 0.5 / (q2 * (q2 - u2));
  const double cw2 = u2 >= 0.0 ? 0.5 / (q2 * (q2 + u2)) : 0.5 * (1.0 - u2 / q2);
  const double e = CN - 2.0 * CS;
  const double g = e - COx;
  const double sD = sqrt(g * g + 4.0 * Kp);
  // X = ((e + COx) - sD) / 2 = 2 (e COx - Kp) / ((e + COx) + sD)
  const double X = (e + COx) > 0.0 ? 2.0 * (e * COx - Kp) / ((e + COx) + sD) : 0.5 * ((e + COx) - sD);
  const double sX = sqrt(X * X + 1e-20);
  const double Xp = X >= 0.0 ? 0.5 * (X + sX) : 0.5 * 1e-20 / (sX - X);
  // H = e - X+ = (e - X) - (X+ - X), e - X = (g + sD) / 2 = 2 Kp / (sD - g) for g < 0
  const double eX = g >= 0.0 ? 0.5 * (g + sD) : 2.0 * Kp / (sD - g);
  const double H = eX - (X >= 0.0 ? 0.5 * 1e-20 / (sX + X) : Xp - X);
  const double sH = sqrt(H * H + 1e-20);
  const double NH3h = H >= 0.0 ? 0.5 * (H + sH) : 0.5 * 1e-20 / (sH - H);
  const double NH3 = w2 * NH3h;
  const double A = w1 * w2 * Xp;
  v[0] = NH3;
  v[1] = CN - NH3;
  v[2] = COx - A;
  v[3] = v[4] = v[5] = A * (1.0 / 3.0);
  if constexpr (DERIV) {
    // columns: d/dCN, d/dCOx, d/dCS
    double am, ap;  // 1 - g / sD and 1 + g / sD
    if (g >= 0.0) {
      am = 4.0 * Kp / (sD * (sD + g));
      ap = 1.0 + g / sD;
    } else {
      am = 1.0 - g / sD;
      ap = 4.0 * Kp / (sD * (sD - g));
    }
    const double dR[3] = {1.0 / den, 0.0, -R / den};
    const double dX[3] = {0.5 * am, 0.5 * ap, -am};
    const double s1 = 0.5 / (q1 * q1 * q1) * inv_width, s2 = 0.5 / (q2 * q2 * q2) * inv_width;
    const double spX = X >= 0.0 ? 0.5 * (1.0 + X / sX) : 0.5 * 1e-20 / (sX * (sX - X));
    const double spH = H >= 0.0 ? 0.5 * (1.0 + H / sH) : 0.5 * 1e-20 / (sH * (sH - H));
  const double cH = H >= 0.0 ? 0.5 * 1e-20 / (sH * (sH + H)) : 0.5 * (1.0 - H / sH);  // 1 - sp'(H)
    // de - dX+ for CN (for CS: -2 times it): 1 - sp'(X) am / 2 = cX + ap/2 - cX ap/2,
    // cX = 1 - sp'(X) (am + ap = 2), which cancels nothing
    const double cX = X >= 0.0 ? 0.5 * 1e-20 / (sX * (sX + X)) : 0.5 * (1.0 - X / sX);
    const double t0 = cX + 0.5 * ap - cX * (0.5 * ap);
    const double dedXp[3] = {t0, -spX * (0.5 * ap), -2.0 * t0};
  // 1 - dNH3/dCN = (1 - w2) + w2 (1 - sp'(H)) + w2 sp'(H) (1 - t0) - dw2/dCN NH3h
  const double one_t0 = (1.0 - cX) * (1.0 - 0.5 * ap);
  const double cdN = cw2 + w2 * cH + w2 * spH * one_t0 - (s2 / den) * NH3h;
#pragma unroll
    for (int c = 0; c < 3; ++c) {
      const double dw1 = s1 * dR[c], dw2 = s2 * dR[c];
      const double dXp = spX * dX[c];
      const double dNH3 = dw2 * NH3h + w2 * (spH * dedXp[c]);
      const double dA = (dw1 * w2 + w1 * dw2) * Xp + w1 * w2 * dXp;
      d[0][c] = dNH3;
      d[1][c] = (c == 0 ? 1.0 : 0.0) - dNH3;
      d[2][c] = (c == 1 ? 1.0 : 0.0) - dA;
      d[3][c] = d[4][c] = d[5][c] = dA * (1.0 / 3.0);
      if (c == 0) {
        d[1][0] = cdN;
        dd[0] = -cdN;
        dd[1] = -dNH3;
      } else if (c == 1) {
        dd[2] = -dA;
        dd[3] = dd[4] = dd[5] = dA * (1.0 / 3.0) - 1.0;
      }
    }
  }
}

}  // namespace mkpp
