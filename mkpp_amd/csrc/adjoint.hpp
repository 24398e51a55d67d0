// Exact discrete adjoint of the wave-per-cell Rosenbrock integrator.
//
// The reference's differentiation templates (solver_variants/
// integrate_fwd_checkpoint.j2:37-295, integrate_adj.j2:50-156) freeze the
// Jacobian and are never run numerically; this is the adjoint of the exact
// discrete map instead.  The forward pass (wave_integrate_cell, MODE 1) keeps
// the state y_n and step size h_n of every accepted step.  Going backwards
// over those steps, one wavefront per cell
//   1. recomputes W = I/(gamma h) - J(y_n), its LU and the stage vectors K_s
//      from the checkpoint (the forward code path, error control dropped);
//   2. runs the stage recursion backwards:
//        Kbar_s = M_s lambda (+ contributions of later stages)
//        rbar_s = W^-T Kbar_s                       (transposed LU solves)
//        Fbar of the stage that evaluated F += rbar_s, Kbar_j += C_sj/h rbar_s
//        at a stage that evaluated F(Y_s): Ybar = J(Y_s)^T Fbar, kbar += dF/dk^T Fbar,
//        ybar += Ybar, Kbar_j += A_sj Ybar
//   3. adds the Jacobian's own dependence on y_n and k:
//        ybar += sum_s d/dy [rbar_s^T J(y_n) K_s]  (second derivatives of the fluxes)
//        kbar += sum_s d/dk [rbar_s^T J(y_n) K_s]
//      and lambda_n = ybar.
// Step sizes are frozen (the accepted sequence of the forward run), so this
// is the gradient of the map that a replay of those steps computes.
//
// Fluxes are mass-action monomials r_g = k_g prod_e y_e^q_e (at most three
// entries), so J^T u and the second-derivative terms are sums over (group,
// entry) partial slots, scattered to species through the generated SP lists.
// Fixed species take part in the fluxes but not in the state: their adjoint
// accumulates over all steps.  Rate constants are per flux group.
#pragma once
#include <type_traits>

#include "wave_engine.hpp"

namespace mkpp {

// The second-derivative pass keeps nu_g . rbar_s for every stage and group,
// (S + 1) x (NG + 1) doubles.  It reuses the LU block, which is free by then,
// when that block is large enough (saprc99: 2015 >= 1055); a small mechanism's
// LU block is not, and the pass gets an area of its own behind the layout.
template <class W, class Tab>
constexpr int adjoint_wh_doubles() {
  return (Tab::S + 1) * (W::NG + 1);
}
template <class W, class Tab>
constexpr bool adjoint_wh_in_lu() {
  return adjoint_wh_doubles<W, Tab>() <= W::SPAN + 1;
}

// The scatter's chunk partials (SZERO + 1 doubles) use the F chunk partials'
// area and the u vector behind it, both dead by then; `extra` pads the two
// where they are shorter.
template <class W>
constexpr int adjoint_cp_extra() {
  constexpr int need = WaveTables<W>::Code::SZERO + 1 - 2 * (W::NV + 1);
  return need > 0 ? need : 0;
}
// The group dots of a J^T pass sit in the partials' area, read before any
// partial is written, where that area holds them.
template <class W>
constexpr bool adjoint_wj_in_scr() {
  return W::NG + 1 <= W::NSCR;
}

// LDS doubles of the adjoint: the forward layout, the scatter padding, the u
// vector of a J^T pass (in the second-derivative pass: Fbar of stage 0), one
// by-row vector per stage that holds Kbar_s until rbar_s replaces it, K of
// every stage by row, the fixed-species adjoint, and the group dots and
// second-derivative areas where the partials' and LU blocks cannot hold them.
template <class W, class Tab>
constexpr int adjoint_lds_doubles() {
  return W::LDS_DOUBLES + adjoint_cp_extra<W>() + (W::NV + 1) + 2 * Tab::S * (W::NV + 1) + (W::NF + 1) +
         (adjoint_wj_in_scr<W>() ? 0 : W::NG + 1) + (adjoint_wh_in_lu<W, Tab>() ? 0 : adjoint_wh_doubles<W, Tab>());
}

template <class W, class Tab>
__device__ __forceinline__ void wave_adjoint_cell(const DeviceArgs& a, const AdjArgs& g, long cell, double* lds,
                                                  int lane) {
  using T_ = WaveTables<W>;
  using C_ = typename T_::Code;
  constexpr int NV = W::NV, NB = W::NB, NG = W::NG, NGP = (NG + 63) / 64, S = Tab::S, NY = W::NY, NF = W::NF;
  double* lu = lds;
  double* ylds = lu + (W::SPAN + 1);
  double* scr = ylds + NY;
  double* facc = scr + W::NSCR;      // [NV+1]    F chunk partials (recomputed stages only)
  double* uv = facc + (NV + 1) + adjoint_cp_extra<W>();  // [NV+1] u of a J^T pass by row
  // [S][NV+1] by row: Kbar_s while the recursion accumulates it, then rbar_s
  // (the transposed solve of stage s reads Kbar_s last).  With uv in front,
  // uv and these rows are the S + 1 vectors of the second-derivative pass
  // (uv holding Fbar of stage 0 there).
  double* rbv = uv + (NV + 1);
  double* kbv = rbv;
  double* kv = rbv + S * (NV + 1);   // [S][NV+1] K_s by row, slot NV = 0.0
  double* fbar = kv + S * (NV + 1);  // [NF]      adjoint of the fixed species
  // scatter chunk partials over facc, the padding and uv (all dead by then)
  double* sacc = facc;
  // [NG+1] nu_g . u of a J^T pass
  double* wj = adjoint_wj_in_scr<W>() ? scr : fbar + (NF + 1);
  // [S+1][NG+1] nu_g . (uv, rbv rows) (second-derivative pass, after the LU is done)
  double* wh = adjoint_wh_in_lu<W, Tab>() ? lu : fbar + (NF + 1) + (adjoint_wj_in_scr<W>() ? 0 : NG + 1);

  // lane slot (b, lane) holds row ROWOF[b * 64 + lane] (255: none)
  int rb[NB], rowv[NB];
  bool valid[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int row = T_::ROWOF[b * 64 + lane];
    valid[b] = row != 255;
    rowv[b] = valid[b] ? row : 0;
    rb[b] = valid[b] ? (int)T_::R[row] : 0;
  }
  const int nst = g.nsteps[cell];
  // A step count outside [0, maxsteps] cannot index the cell's checkpoint
  // slice: the cell's outputs are NaN and no checkpoint is read.
  if (nst < 0 || nst > g.maxsteps) {
    const double bad = __builtin_nan("");
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if (valid[b]) g.lambda[cell + a.stride * (long)T_::VARPHYS[rowv[b]]] = bad;
    if (lane < NF) g.lambda[cell + a.stride * (long)T_::FIXPHYS[lane]] = bad;
    for (int gi = lane; gi < NG; gi += 64) g.kbar[cell + g.kstride * (long)gi] = bad;
    return;
  }

  // ---- environment, fixed species, rate constants (as the forward pass) ----
  if (lane < NF) ylds[NV + lane] = a.conc[cell + a.stride * (long)T_::FIXPHYS[lane]];
  if (lane == 0) ylds[W::ONE] = 1.0;
  const double Tk = a.temp ? a.temp[cell] : a.temp_default;
  const double rho = a.air_density ? a.air_density[cell] : a.rho_default;
  const double Mtb = (W::THIRD_BODY >= 0) ? a.conc[cell + a.stride * (long)(W::THIRD_BODY >= 0 ? W::THIRD_BODY : 0)] : rho;
  double kreg[NGP], kbar[NGP];
#pragma unroll
  for (int p = 0; p < NGP; ++p) {
    const int gi = p * 64 + lane;
    kreg[p] = 0.0;
    kbar[p] = 0.0;
    if (gi < NG) {
      if (g.k != nullptr) {
        kreg[p] = g.k[cell + g.kstride * (long)gi];
      } else {
        const int law = T_::LAW[gi];
        const int ph = T_::PHOTO[gi];
        const double jv = (law == 0 && a.jvals != nullptr) ? a.jvals[cell * a.jcstride + a.jstride * ph] : 0.0;
        kreg[p] = wave_rate(law, &T_::PARAMS[gi * 7], Tk, Mtb, jv);
      }
    }
  }

  // ---- lambda = dL/dy(T) ----
  double lam[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) lam[b] = valid[b] ? g.lambda[cell + a.stride * (long)T_::VARPHYS[rowv[b]]] : 0.0;
  if (lane < NF) fbar[lane] = g.lambda[cell + a.stride * (long)T_::FIXPHYS[lane]];
  for (int i = lane; i < S * (NV + 1); i += 64) kv[i] = 0.0;
  // slot NV of the by-row vectors stays 0.0: the padding terms of the gather
  // read it (uv's is set again before each use: the scatter may overwrite it)
  if (lane < S) rbv[lane * (NV + 1) + NV] = 0.0;

  auto store_y = [&](const double (&y)[NB]) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if (valid[b]) ylds[rowv[b]] = y[b];
  };

  auto eval_rates = [&](double (&f)[NB]) { wave_eval_f<W>(ylds, scr, kreg, valid, f, lane); };

  // reactant powers of group gi at the state in ylds: pw = y^q, pm1 = q y^(q-1),
  // pm2 = q (q-1) y^(q-2) (zero where the power is too small)
  auto monomial = [&](int gi, double (&pw)[3], double (&pm1)[3], double (&pm2)[3]) {
    const uint8_t* s = &T_::GPOWS[gi * 3];
    const uint8_t* q = &T_::GPOWP[gi * 3];
#pragma unroll
    for (int e = 0; e < 3; ++e) {
      const double y = ylds[s[e]];
      const double y2 = y * y;
      const int qe = q[e];
      pw[e] = qe == 0 ? 1.0 : (qe == 1 ? y : (qe == 2 ? y2 : y2 * y));
      pm1[e] = qe == 0 ? 0.0 : (qe == 1 ? 1.0 : (qe == 2 ? 2.0 * y : 3.0 * y2));
      pm2[e] = qe <= 1 ? 0.0 : (qe == 2 ? 2.0 : 6.0 * y);
    }
  };

  auto assemble_w = [&](double inv_g_dt) {
#pragma unroll
    for (int p = 0; p < NGP; ++p) {
      const int gi = p * 64 + lane;
      if (gi < NG) {
        const int np = T_::GNPOW[gi];
        const uint8_t* s = &T_::GPOWS[gi * 3];
        const uint8_t* q = &T_::GPOWP[gi * 3];
        const int off = T_::POFF[gi];
        double yv[3], pw[3], pm1[3];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          yv[e] = ylds[s[e]];
          const double y2 = yv[e] * yv[e];
          pw[e] = q[e] == 0 ? 1.0 : (q[e] == 1 ? yv[e] : (q[e] == 2 ? y2 : y2 * yv[e]));
          pm1[e] = q[e] <= 1 ? 1.0 : (q[e] == 2 ? yv[e] : y2);
        }
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          double t = (kreg[p] * (double)q[e]) * pm1[e];
#pragma unroll
          for (int o = 0; o < 3; ++o)
            if (o != e) t = t * pw[o];
          if (e < np) scr[off + e] = t;
        }
      }
    }
    if (lane == 0) scr[W::NPART] = 0.0;
    wave_zero_lu<W>(lu, lane);
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if (valid[b]) lu[rb[b] + rowv[b]] = inv_g_dt;
#pragma unroll
    for (int p = 0; p < C_::NWP; ++p) {
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < C_::TW; ++q) {
        const uint32_t tm = T_::CW_TERM[(p * C_::TW + q) * 64 + lane];
        acc = fma(T_::COEF[tm >> 16], scr[tm & 0xFFFFu], acc);
      }
      // every lane updates: padding lanes subtract 0 from an entry no other
      // lane of the pass touches
      const int ad = T_::CW_ADDR[p * 64 + lane];
      lu[ad] = lu[ad] - acc;
    }
  };

  // out[s * (NG + 1) + g] = nu_g . u_s for every group g and the NVEC vectors
  // u_s = u + s * (NV + 1) (by row).  Fixed-trip chunks (CG tables): one chunk of
  // a group's (row, coefficient) terms per lane and pass.
  auto group_dots = [&](auto nvec_c, const double* u, double* out) {
    constexpr int NVEC = decltype(nvec_c)::value;
    for (int i = lane; i < NVEC * (NG + 1); i += 64) out[i] = 0.0;
#pragma unroll
    for (int p = 0; p < C_::NGAT; ++p) {
      double acc[NVEC];
#pragma unroll
      for (int s = 0; s < NVEC; ++s) acc[s] = 0.0;
#pragma unroll
      for (int q = 0; q < C_::TG; ++q) {
        const uint32_t tm = T_::CG_TERM[(p * C_::TG + q) * 64 + lane];
        const double cf = T_::COEF[tm >> 8];
        const int r = tm & 255u;
#pragma unroll
        for (int s = 0; s < NVEC; ++s) acc[s] = fma(cf, u[s * (NV + 1) + r], acc[s]);
      }
      const int gg = T_::CG_GRP[p * 64 + lane];
      if (gg != 0xFFFF) {
#pragma unroll
        for (int s = 0; s < NVEC; ++s) out[s * (NG + 1) + gg] += acc[s];
      }
    }
  };

  // scr[partial slot] summed per species slot: rows into out[], fixed species
  // into fbar.  Two levels (CS tables): chunks of up to TS partial slots, any
  // species, summed by lane into cp[chunk]; then each species' lane adds its
  // chunks in order.  cp spans the F chunk partials, the padding and the u
  // vector, whose contents are consumed before the scatter.
  auto scatter = [&](double (&out)[NB]) {
    double* cp = sacc;
    if (lane == 0) scr[W::NPART] = 0.0;
#pragma unroll
    for (int p = 0; p < C_::NSP; ++p) {
      const uint32_t* tw = &T_::CS_TERM[(p * 64 + lane) * (C_::TS / 2)];
      double acc = 0.0;
#pragma unroll
      for (int q = 0; q < C_::TS / 2; ++q) {
        const uint32_t two = tw[q];
        acc += scr[two & 0xFFFFu];
        acc += scr[two >> 16];
      }
      cp[p * 64 + lane] = acc;
    }
    if (lane == 0) cp[C_::SZERO] = 0.0;
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const uint16_t* pp = &T_::CS_PART[(b * 64 + lane) * C_::SMAXCP];
      double sum = cp[pp[0]];
#pragma unroll
      for (int c = 1; c < C_::SMAXC[b]; ++c) sum += cp[pp[c]];
      out[b] = valid[b] ? sum : 0.0;
    }
    if (lane < NF) {
      const uint16_t* pp = &T_::CS_FIX[lane * C_::SFMAXCP];
      double sum = cp[pp[0]];
#pragma unroll
      for (int c = 1; c < C_::SFMAXC; ++c) sum += cp[pp[c]];
      fbar[lane] += sum;
    }
  };

  // J(Y)^T u at the state in ylds (u by row in uv), plus kbar += u . dF/dk
  // The group-block loops of the two passes stay rolled (four copies of the
  // second-derivative code would not fit the register file); the block's rate
  // constant and kbar accumulator are selected by the runtime block index.
  auto kreg_at = [&](int p) {
    double v = 0.0;
#pragma unroll
    for (int q = 0; q < NGP; ++q)
      if (q == p) v = kreg[q];
    return v;
  };
  auto kbar_add = [&](int p, double v) {
#pragma unroll
    for (int q = 0; q < NGP; ++q)
      if (q == p) kbar[q] += v;
  };

  auto jt_pass = [&](double (&out)[NB]) {
    if (lane == 0) uv[NV] = 0.0;
    group_dots(std::integral_constant<int, 1>{}, uv, wj);
    // every dot is read before the first partial is written (wj may be the
    // partials' area; one wavefront's LDS accesses complete in order)
    double wg[NGP];
#pragma unroll
    for (int p = 0; p < NGP; ++p) wg[p] = p * 64 + lane < NG ? wj[p * 64 + lane] : 0.0;
#pragma unroll
    for (int p = 0; p < NGP; ++p) {
      const int gi = p * 64 + lane;
      if (gi < NG) {
        const double w = wg[p];
        double pw[3], pm1[3], pm2[3];
        monomial(gi, pw, pm1, pm2);
        kbar_add(p, w * ((pw[0] * pw[1]) * pw[2]));
        const double kp = kreg_at(p);
        const int np = T_::GNPOW[gi];
        const int off = T_::POFF[gi];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          double d = pm1[e];
#pragma unroll
          for (int o = 0; o < 3; ++o)
            if (o != e) d *= pw[o];
          if (e < np) scr[off + e] = (w * kp) * d;
        }
      }
    }
    scatter(out);
  };

  // second-derivative terms at y_n: sum_s d/dy [rbar_s^T J K_s] and d/dk
  // Also carries the J^T pass of stage 0, which is evaluated at y_n too: its
  // Fbar is in uv, the vector in front of the rbv rows, and its terms join
  // the same partial slots.
  auto hess_pass = [&](double (&out)[NB]) {
    if (lane == 0) uv[NV] = 0.0;
    group_dots(std::integral_constant<int, S + 1>{}, uv, wh);
#pragma unroll 1
    for (int p = 0; p < NGP; ++p) {
      const int gi = p * 64 + lane;
      if (gi < NG) {
        double pw[3], pm1[3], pm2[3];
        monomial(gi, pw, pm1, pm2);
        // K by the entry's own row: a unit-value species reads 1.0 in the
        // monomial (as W is assembled) but its Jacobian column multiplies its
        // stage increment
        const uint8_t* sk = &T_::GPOWK[gi * 3];
        // first and second derivatives of the monomial (without k)
        double d1[3], d2[3][3];
#pragma unroll
        for (int e = 0; e < 3; ++e) {
          double d = pm1[e];
#pragma unroll
          for (int o = 0; o < 3; ++o)
            if (o != e) d *= pw[o];
          d1[e] = d;
#pragma unroll
          for (int f = 0; f < 3; ++f) {
            double h;
            if (f == e) {
              h = pm2[e];
#pragma unroll
              for (int o = 0; o < 3; ++o)
                if (o != e) h *= pw[o];
            } else {
              h = pm1[e] * pm1[f];
#pragma unroll
              for (int o = 0; o < 3; ++o)
                if (o != e && o != f) h *= pw[o];
            }
            d2[e][f] = h;
          }
        }
        double ws[S];
#pragma unroll
        for (int s = 0; s < S; ++s) ws[s] = wh[(s + 1) * (NG + 1) + gi];
        const double w0 = wh[gi];
        double val[3] = {w0 * d1[0], w0 * d1[1], w0 * d1[2]};
        double kb = w0 * ((pw[0] * pw[1]) * pw[2]);
#pragma unroll
        for (int s = 0; s < S; ++s) {
          const double w = ws[s];
          double kvs[3];
#pragma unroll
          for (int e = 0; e < 3; ++e) kvs[e] = kv[s * (NV + 1) + (sk[e] < NV ? sk[e] : NV)];
          double dk = 0.0;
#pragma unroll
          for (int e = 0; e < 3; ++e) {
            dk = fma(kvs[e], d1[e], dk);
            double de = 0.0;
#pragma unroll
            for (int f = 0; f < 3; ++f) de = fma(kvs[f], d2[f][e], de);
            val[e] = fma(w, de, val[e]);
          }
          kb = fma(w, dk, kb);
        }
        kbar_add(p, kb);
        const double kp = kreg_at(p);
        const int np = T_::GNPOW[gi];
        const int off = T_::POFF[gi];
#pragma unroll
        for (int e = 0; e < 3; ++e)
          if (e < np) scr[off + e] = kp * val[e];
      }
    }
    scatter(out);
  };

  for (int n = nst - 1; n >= 0; --n) {
    const double dt = g.hs[cell * (long)g.maxsteps + n];
    const double* ck = g.ckpt + (cell * (long)g.maxsteps + n) * NV;
    double x[NB];
#pragma unroll
    for (int b = 0; b < NB; ++b) x[b] = valid[b] ? ck[rowv[b]] : 0.0;
    const double inv_g_dt = 1.0 / (Tab::G * dt);
    int ln = lane;
    asm volatile("" : "+v"(ln));
    double K[S][NB], F[NB], rdiag[NB];
    store_y(x);
    assemble_w(inv_g_dt);
    C_::factor(lu, rb, ln);
#pragma unroll
    for (int b = 0; b < NB; ++b) rdiag[b] = valid[b] ? 1.0 / lu[rb[b] + rowv[b]] : 0.0;

    // ---- forward stages, recomputed from the checkpoint ----
#pragma unroll 1
    for (int s = 0; s < S; ++s) {
      double wv[NB];
      const int base = s * (s - 1) / 2;
      if (s == 0 || Tab::NEWF[s]) {
        if (s > 0) {
          double Y[NB];
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            double v = x[b];
#pragma unroll
            for (int j = 0; j < S - 1; ++j) {
              const double c = j < s ? Tab::A[base + j] : 0.0;
              if (c == 0.0) continue;
              v = (c == 1.0) ? v + K[j][b] : v + c * K[j][b];
            }
            Y[b] = v;
          }
          store_y(Y);
        }
        eval_rates(F);
      }
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        double v = F[b];
#pragma unroll
        for (int j = 0; j < S - 1; ++j) {
          const double c = j < s ? Tab::C[base + j] : 0.0;
          if (c == 0.0) continue;
          v = v + (c / dt) * K[j][b];
        }
        wv[b] = valid[b] ? v : 0.0;
      }
      int ls = ln;
      asm volatile("" : "+v"(ls));
      double xs[NB];
      C_::forward(lu, rb, wv, ls);
      C_::backward(lu, rb, wv, rdiag, xs, ls);
#pragma unroll
      for (int j = 0; j < S; ++j)
        if (j == s) {
#pragma unroll
          for (int b = 0; b < NB; ++b) K[j][b] = xs[b];
        }
    }
    // K_s by row for the recursion and the second-derivative pass
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int b = 0; b < NB; ++b)
        if (valid[b]) kv[s * (NV + 1) + rowv[b]] = K[s][b];

    // ---- backward stage recursion (K and Kbar live in LDS, by row) ----
    double yb[NB], pend[NB];
#pragma unroll
    for (int s = 0; s < S; ++s)
#pragma unroll
      for (int b = 0; b < NB; ++b)
        if (valid[b]) kbv[s * (NV + 1) + rowv[b]] = Tab::M[s] * lam[b];
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      yb[b] = lam[b];
      pend[b] = 0.0;
    }
#pragma unroll 1
    for (int s = S - 1; s >= 0; --s) {
      const int base = s * (s - 1) / 2;
      double kbs[NB], z[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) kbs[b] = valid[b] ? kbv[s * (NV + 1) + rowv[b]] : 0.0;
      int ls = ln;
      asm volatile("" : "+v"(ls));
      C_::tforward(lu, rowv, kbs, rdiag, z, ls);
      C_::tbackward(lu, rowv, z, ls);
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        if (!valid[b]) z[b] = 0.0;
        if (valid[b]) rbv[s * (NV + 1) + rowv[b]] = z[b];
        pend[b] += z[b];
      }
      for (int j = 0; j < s; ++j) {
        const double c = Tab::C[base + j];
        if (c == 0.0) continue;
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (valid[b]) {
            double* kb = &kbv[j * (NV + 1) + rowv[b]];
            *kb = fma(c / dt, z[b], *kb);
          }
      }
      if (s == 0 || Tab::NEWF[s]) {
        double Y[NB];
#pragma unroll
        for (int b = 0; b < NB; ++b) Y[b] = x[b];
        for (int j = 0; j < s; ++j) {
          const double c = Tab::A[base + j];
          if (c == 0.0) continue;
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            const double kj = valid[b] ? kv[j * (NV + 1) + rowv[b]] : 0.0;
            Y[b] = (c == 1.0) ? Y[b] + kj : Y[b] + c * kj;
          }
        }
        store_y(Y);
        if (s == 0) {
          // stage 0 (Y = y_n): its J^T pass is merged into the second-derivative pass
#pragma unroll
          for (int b = 0; b < NB; ++b)
            if (valid[b]) uv[rowv[b]] = pend[b];
          continue;
        }
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (valid[b]) uv[rowv[b]] = pend[b];
        double Yb[NB];
        jt_pass(Yb);
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          yb[b] += Yb[b];
          pend[b] = 0.0;
        }
        for (int j = 0; j < s; ++j) {
          const double c = Tab::A[base + j];
          if (c == 0.0) continue;
#pragma unroll
          for (int b = 0; b < NB; ++b)
            if (valid[b]) {
              double* kb = &kbv[j * (NV + 1) + rowv[b]];
              *kb = fma(c, Yb[b], *kb);
            }
        }
      }
    }
    // ---- the Jacobian's dependence on y_n (ylds holds y_n after stage 0) ----
    double Hb[NB];
    hess_pass(Hb);
#pragma unroll
    for (int b = 0; b < NB; ++b) lam[b] = yb[b] + Hb[b];
  }

  // ---- outputs ----
#pragma unroll
  for (int b = 0; b < NB; ++b)
    if (valid[b]) g.lambda[cell + a.stride * (long)T_::VARPHYS[rowv[b]]] = lam[b];
  if (lane < NF) g.lambda[cell + a.stride * (long)T_::FIXPHYS[lane]] = fbar[lane];
#pragma unroll
  for (int p = 0; p < NGP; ++p)
    if (p * 64 + lane < NG) g.kbar[cell + g.kstride * (long)(p * 64 + lane)] = kbar[p];
}

// One cell per wavefront at the XCD-aware cell order of the forward kernel.
template <class W, class Tab>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) wave_adjoint_kernel(DeviceArgs a, AdjArgs g, long base) {
  extern __shared__ double lds[];
  const long cell = wave_cell_of_block(base);
  if (cell < a.ncells) wave_adjoint_cell<W, Tab>(a, g, cell, lds, threadIdx.x);
}

template <class W, class Tab, int MODE>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(2, 2))) wave_record_kernel(DeviceArgs a, AdjArgs g, long base) {
  extern __shared__ double lds[];
  const long cell = wave_cell_of_block(base);
  if (cell < a.ncells) wave_integrate_cell<W, Tab, MODE>(a, cell, lds, threadIdx.x, &g);
}

// Tangent-linear replay (MODE 4): more registers than the forward kernel (the
// tangent stage vectors), so no occupancy target is forced on it.
template <class W, class Tab>
__global__ void __launch_bounds__(64) wave_tlm_kernel(DeviceArgs a, AdjArgs g, long base) {
  extern __shared__ double lds[];
  const long cell = wave_cell_of_block(base);
  if (cell < a.ncells) wave_integrate_cell<W, Tab, 4>(a, cell, lds, threadIdx.x, &g);
}

// Rate constants per flux group and cell (the values the laws give; the
// differentiable inputs of the adjoint).
template <class W>
__global__ void __launch_bounds__(64) wave_rate_constants_kernel(DeviceArgs a, double* k, long kstride) {
  using T_ = WaveTables<W>;
  const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= a.ncells) return;
  const double Tk = a.temp ? a.temp[cell] : a.temp_default;
  const double rho = a.air_density ? a.air_density[cell] : a.rho_default;
  const double Mtb = (W::THIRD_BODY >= 0) ? a.conc[cell + a.stride * (long)(W::THIRD_BODY >= 0 ? W::THIRD_BODY : 0)] : rho;
  for (int gi = 0; gi < W::NG; ++gi) {
    const int law = T_::LAW[gi];
    const int ph = T_::PHOTO[gi];
    const double jv = (law == 0 && a.jvals != nullptr) ? a.jvals[cell * a.jcstride + a.jstride * ph] : 0.0;
    k[cell + kstride * (long)gi] = wave_rate(law, &T_::PARAMS[gi * 7], Tk, Mtb, jv);
  }
}

// mode 0: rate constants only; 1: checkpointed forward; 2: replay; 3: adjoint;
// 4: tangent-linear replay; every tableau.
template <class W>
int launch_wave_diff(int mode, const DeviceArgs& a, const AdjArgs& g, hipStream_t s) {
  if (a.ncells <= 0) return 0;
  if (mode == 0) {
    wave_rate_constants_kernel<W><<<(unsigned)((a.ncells + 63) / 64), 64, 0, s>>>(a, g.kbar, g.kstride);
    return hipGetLastError() == hipSuccess ? 0 : -2;
  }
  const size_t fwd_lds = sizeof(double) * W::LDS_DOUBLES;
#define MKPP_DIFF_CASE(TAB)                                                                              \
  if (mode == 1) {                                                                                       \
    wave_allow_lds(&wave_record_kernel<W, TAB, 1>, fwd_lds);                                             \
    wave_record_kernel<W, TAB, 1><<<grid, 64, fwd_lds, s>>>(a, g, base);                                 \
  } else if (mode == 2) {                                                                                \
    wave_allow_lds(&wave_record_kernel<W, TAB, 2>, fwd_lds);                                             \
    wave_record_kernel<W, TAB, 2><<<grid, 64, fwd_lds, s>>>(a, g, base);                                 \
  } else if (mode == 4) {                                                                                \
    const size_t tl = sizeof(double) * wave_tlm_lds_doubles<W>();                                        \
    wave_allow_lds(&wave_tlm_kernel<W, TAB>, tl);                                                        \
    wave_tlm_kernel<W, TAB><<<grid, 64, tl, s>>>(a, g, base);                                            \
  } else {                                                                                               \
    const size_t al = sizeof(double) * adjoint_lds_doubles<W, TAB>();                                    \
    wave_allow_lds(&wave_adjoint_kernel<W, TAB>, al);                                                    \
    wave_adjoint_kernel<W, TAB><<<grid, 64, al, s>>>(a, g, base);                                        \
  }
  if (a.method < ROS2 || a.method > RODAS4) return -1;
  for (long base = 0; base < a.ncells; base += WAVE_CHUNK_CELLS) {
    const unsigned grid = wave_chunk_grid(a.ncells, base);
    switch (a.method) {
      case ROS2: MKPP_DIFF_CASE(TabRos2) break;
      case ROS3: MKPP_DIFF_CASE(TabRos3) break;
      case ROS4: MKPP_DIFF_CASE(TabRos4) break;
      case RODAS3: MKPP_DIFF_CASE(TabRodas3) break;
      default: MKPP_DIFF_CASE(TabRodas4) break;
    }
  }
#undef MKPP_DIFF_CASE
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace mkpp
