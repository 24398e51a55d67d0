// Wave-per-cell Rosenbrock engine for mechanisms too large for one lane
// (saprc99: 78 variable species, 1547 LU entries, 9654 products per factor).
//
// One 64-lane wavefront integrates one cell (one cell per workgroup, so the
// hardware dispatcher balances uneven cells).  LDS holds the W/LU matrix in row-displacement
// storage (element (i, j) at R[i] + j), the stage state y for F gathers and a
// scratch area for fluxes / Jacobian partials.  Stage vectors live in
// registers, lane = row (bank b covers rows 64b .. 64b+63).
//
// Arithmetic follows the oracle's order per entry: F_i and J_ij accumulate in
// flux-group order, every LU entry subtracts its products in ascending pivot
// order then divides by the pivot (right-looking schedule, same sequence per
// entry as the reference's Crout-ordered plan), the error norm sums rows in
// logical order.  Two differences, both within rounding: LU, solve and
// accumulation updates use fused multiply-add, and the backward substitution
// subtracts a row's products in descending column order.
#pragma once
#include <hip/hip_runtime.h>

#include "mkpp_device.hpp"
#include "runtime.hpp"

namespace mkpp {

__device__ __forceinline__ double readlane_d(double v, int lane) {
  const int lo = __builtin_amdgcn_readlane(__double2loint(v), lane);
  const int hi = __builtin_amdgcn_readlane(__double2hiint(v), lane);
  return __hiloint2double(hi, lo);
}

// Constant lane masks of the generated straight-line code, as SGPR pairs that
// v_cndmask and s_and_saveexec take directly (no per-lane shift and compare).
// The two moves sit in an asm statement whose input is an SGPR the caller
// launders once per call, so they stay next to their use inside the step loop
// rather than being hoisted out of it (hundreds of live masks would spill);
// ID makes every site distinct, so equal masks are not merged into one long
// live range.  The statement is not volatile: it orders nothing, and the
// loads around it still run ahead of the substitution chain.
template <uint32_t LO, uint32_t HI, int ID>
__device__ __forceinline__ uint64_t lane_mask(int opaque) {
  uint32_t lo, hi;
  asm("s_mov_b32 %0, %2\n\ts_mov_b32 %1, %3 ; mask %4" : "=s"(lo), "=s"(hi) : "i"(LO), "i"(HI), "i"(ID), "s"(opaque));
  return ((uint64_t)hi << 32) | lo;
}

// v in the lanes of m, +0.0 elsewhere.  The empty asm hides that v is a load,
// so the select is not turned back into a branch around a predicated load.
__device__ __forceinline__ double lane_keep(uint64_t m, double v) {
  asm("" : "+v"(v));
  return __builtin_amdgcn_inverse_ballot_w64(m) ? v : 0.0;
}

// Rate constant of one flux group (src/mkpp/lowering.py:199-313); P holds the
// law's numeric parameters in the order of mkpp_amd.gen.plan.LAW_PARAMS.
__device__ __forceinline__ double wave_rate(int law, const double* P, double T, double M, double jv) {
  switch (law) {
    case 0:  // PHOTOLYSIS
      return jv;
    case 1: {  // ARRHENIUS: A (T/300)^B exp(C/T); unit factors skipped exactly
      double k = P[0];
      k = (P[1] != 0.0) ? k * pow(T / 300.0, P[1]) : k * 1.0;
      if (P[2] != 0.0) k = k * exp(P[2] / T);
      return k;
    }
    case 2: {  // TROE: A0 B0 C0 A1 B1 C1 Fc
      const double k0 = P[0] * exp(P[2] / T) * pow(T / 300.0, P[1]);
      const double ki = P[3] * exp(P[5] / T) * pow(T / 300.0, P[4]);
      const double K0 = k0 * M;
      const double r = K0 / ki;
      const double lg = log10(r);
      const double ex = 1.0 / (1.0 + lg * lg);
      return K0 / (1.0 + r) * pow(P[6], ex);
    }
    case 3: {  // EP2: A0 C0 A2 C2 A3 C3
      const double K0 = P[0] * exp(P[1] / T);
      const double K2 = P[2] * exp(P[3] / T);
      const double K3 = P[4] * exp(P[5] / T) * M;
      return K0 + K3 / (1.0 + K3 / K2);
    }
    case 4: {  // EP3: A1 C1 A2 C2
      const double K1 = P[0] * exp(P[1] / T);
      const double K2 = P[2] * exp(P[3] / T);
      return K1 + K2 * M;
    }
    case 5:  // HETEROGENEOUS: 0.25 gamma (v_gas = S_a = 1)
      return 0.25 * P[0];
    default:  // CONSTANT
      return P[0];
  }
}

template <class W>
struct WaveTables;  // specialised per mechanism by MKPP_WAVE_MECHANISM

// bank-select without dynamic register-array indexing
template <int NB>
__device__ __forceinline__ double bank(const double (&v)[NB], int row) {
  if constexpr (NB == 1) {
    return v[0];
  } else if constexpr (NB == 2) {
    return row < 64 ? v[0] : v[1];
  } else {
    double r = v[0];
#pragma unroll
    for (int b = 1; b < NB; ++b) r = row >= 64 * b ? v[b] : r;
    return r;
  }
}

// Zeroes the LU storage lu[0, SPAN] with 16-byte stores (the LDS block starts
// 16-byte aligned), unrolled: no loop branch per row of 64 lanes.
template <class W>
__device__ __forceinline__ void wave_zero_lu(double* lu, int lane) {
  constexpr int NZ = W::SPAN + 1;
  double2* lu2 = reinterpret_cast<double2*>(lu);
#pragma unroll
  for (int i = 0; i < NZ / 128; ++i) lu2[i * 64 + lane] = make_double2(0.0, 0.0);
#pragma unroll
  for (int i = (NZ / 128) * 128; i < NZ; i += 64)
    if (i + lane < NZ) lu[i + lane] = 0.0;
}

// F(y) by row slot from the state in ylds (rows, fixed species, ONE), in
// three steps over the scratch area:
// 1. lanes take flux groups: scr[g] = k_g * y_a * y_b * y_c, scr[NG] = 0.0;
// 2. each row's terms (coefficient x flux, in the row's order) are cut into
//    chunks of at most TF; lane j % 64 of pass j / 64 sums chunk j into the
//    partial slot scr[PBASE + j] (every lane writes; padding adds 0 x 0.0);
// 3. the lane of each row adds its chunks' partials in chunk order.
// A row thus sums its terms in the reference's order (lowering.py, F_i over
// the flux groups) grouped by chunks, a difference of rounding only.  No
// loop bound depends on data, so every table load and LDS gather of a step
// can be issued at once.
//
// wave_row_sums is steps 2 and 3 alone: any per-group values in scr[0, NG)
// (with scr[NG] = 0.0) become sum_g nu_ig * scr[g] per row, the F sums'
// order and grouping.  The tangent-linear mode uses it for J v and (dJ) K.
template <class W, int NB>
__device__ __forceinline__ void wave_row_sums(double* __restrict__ scr, const bool (&valid)[NB], double (&f)[NB],
                                              int lane) {
  using T_ = WaveTables<W>;
  using C_ = typename T_::Code;
  static_assert(C_::PBASE + C_::NFP * 64 <= W::NSCR + W::NV + 1, "F partial slots exceed the scratch and row-sum areas");
#pragma unroll
  for (int p = 0; p < C_::NFP; ++p) {
    const uint32_t* tw = &T_::CF_TERM[(p * 64 + lane) * (C_::TF / 2)];
    double acc = 0.0;
#pragma unroll
    for (int q = 0; q < C_::TF / 2; ++q) {
      // two 16-bit terms, flux slot in the low FBITS bits, coefficient above
      constexpr uint32_t FB = C_::FBITS, FM = (1u << FB) - 1u;
      const uint32_t two = tw[q];
      acc = fma(T_::COEF[(two & 0xFFFFu) >> FB], scr[two & FM], acc);
      acc = fma(T_::COEF[two >> (16 + FB)], scr[(two >> 16) & FM], acc);
    }
    scr[C_::PBASE + p * 64 + lane] = acc;
  }
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const uint16_t* pp = &T_::CF_PART[(b * 64 + lane) * C_::FMAXCP];
    double sum = scr[pp[0]];
#pragma unroll
    for (int c = 1; c < C_::FMAXC[b]; ++c) sum = sum + scr[pp[c]];
    f[b] = valid[b] ? sum : 0.0;
  }
}

template <class W, int NB, int NGP>
__device__ __forceinline__ void wave_eval_f(const double* __restrict__ ylds, double* __restrict__ scr,
                                            const double (&kreg)[NGP], const bool (&valid)[NB], double (&f)[NB],
                                            int lane) {
  using T_ = WaveTables<W>;
  constexpr int NG = W::NG;
#pragma unroll
  for (int p = 0; p < NGP; ++p) {
    const int g = p * 64 + lane;
    if (g < NG) {
      // ((1.0 * y_0) * y_1) * ... over the NFAC factor slots (padding reads 1.0)
      const uint8_t* fa = &T_::GFAC[g * W::NFAC];
      double m = 1.0 * ylds[fa[0]];
#pragma unroll
      for (int e = 1; e < W::NFAC; ++e) m = m * ylds[fa[e]];
      scr[g] = kreg[p] * m;
    }
  }
  if (lane == 0) scr[NG] = 0.0;
  wave_row_sums<W>(scr, valid, f, lane);
}

// Tangent-linear group values (MODE 4 of wave_integrate_cell).  Fluxes are
// mass-action monomials r_g = k_g prod_e y_e^q_e over at most three entries
// (distinct reactants; powers up to 3).
// Derivatives of the monomial at the state y (by y slot): d1[e] = d/dy_e,
// d2[e][f] = d2/dy_e dy_f (the power rule on each entry).
template <class W>
__device__ __forceinline__ void wave_monomial_derivs(const double* __restrict__ y, int g, double& mono,
                                                     double (&d1)[3], double (&d2)[3][3]) {
  using T_ = WaveTables<W>;
  const uint8_t* sl = &T_::GPOWS[g * 3];
  const uint8_t* ql = &T_::GPOWP[g * 3];
  double pw[3], pm1[3], pm2[3];
#pragma unroll
  for (int e = 0; e < 3; ++e) {
    const double v = y[sl[e]];
    const double v2 = v * v;
    const int q = ql[e];
    pw[e] = q == 0 ? 1.0 : (q == 1 ? v : (q == 2 ? v2 : v2 * v));
    pm1[e] = q == 0 ? 0.0 : (q == 1 ? 1.0 : (q == 2 ? 2.0 * v : 3.0 * v2));
    pm2[e] = q <= 1 ? 0.0 : (q == 2 ? 2.0 : 6.0 * v);
  }
  mono = (pw[0] * pw[1]) * pw[2];
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
}

// scr[g] = dr_g = dk_g mono_g(Y) + k_g sum_e d1_e(Y) dY_e: with wave_row_sums,
// J(Y) dY + F_k(Y) dk.  y, dy by y slot (dy of ONE is 0).
template <class W, int NGP>
__device__ __forceinline__ void wave_tlm_group_jvp(const double* __restrict__ y, const double* __restrict__ dy,
                                                   double* __restrict__ scr, const double (&kreg)[NGP],
                                                   const double (&dkreg)[NGP], int lane) {
  using T_ = WaveTables<W>;
  constexpr int NG = W::NG;
#pragma unroll
  for (int p = 0; p < NGP; ++p) {
    const int g = p * 64 + lane;
    if (g < NG) {
      double mono, d1[3], d2[3][3];
      wave_monomial_derivs<W>(y, g, mono, d1, d2);
      const uint8_t* sl = &T_::GPOWS[g * 3];
      double jv = 0.0;
#pragma unroll
      for (int e = 0; e < 3; ++e) jv = fma(d1[e], dy[sl[e]], jv);
      scr[g] = fma(dkreg[p], mono, kreg[p] * jv);
    }
  }
  if (lane == 0) scr[NG] = 0.0;
}

// scr[g] = k_g sum_e sum_f d2_ef(y) K_e dy_f + dk_g sum_e d1_e(y) K_e: with
// wave_row_sums, the derivative of J(y) K along (dy, dk).  y, dy, K by y slot
// (K of fixed species and ONE is 0).  The monomial is evaluated as W is
// assembled, a unit-value species reading 1.0 (slot ONE); its Jacobian column
// still multiplies its own stage increment, so K is read through GPOWK (the
// species' row) rather than GPOWS.
template <class W, int NGP>
__device__ __forceinline__ void wave_tlm_group_djk(const double* __restrict__ y, const double* __restrict__ dy,
                                                   const double* __restrict__ kv, double* __restrict__ scr,
                                                   const double (&kreg)[NGP], const double (&dkreg)[NGP], int lane) {
  using T_ = WaveTables<W>;
  constexpr int NG = W::NG;
#pragma unroll
  for (int p = 0; p < NGP; ++p) {
    const int g = p * 64 + lane;
    if (g < NG) {
      double mono, d1[3], d2[3][3];
      wave_monomial_derivs<W>(y, g, mono, d1, d2);
      const uint8_t* sl = &T_::GPOWS[g * 3];
      const uint8_t* sk = &T_::GPOWK[g * 3];
      double kk[3], dd[3];
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        kk[e] = kv[sk[e]];
        dd[e] = dy[sl[e]];
      }
      double h = 0.0, jk = 0.0;
#pragma unroll
      for (int e = 0; e < 3; ++e) {
        jk = fma(d1[e], kk[e], jk);
#pragma unroll
        for (int f = 0; f < 3; ++f) h = fma(d2[e][f] * kk[e], dd[f], h);
      }
      scr[g] = fma(dkreg[p], jk, kreg[p] * h);
    }
  }
  if (lane == 0) scr[NG] = 0.0;
}

// LDS doubles of the tangent-linear mode: the forward layout plus y_n, dy_n,
// dY_s and K_s by y slot.
template <class W>
constexpr int wave_tlm_lds_doubles() {
  return W::LDS_DOUBLES + 4 * W::NY;
}

// ---------------------------------------------------------------------------
// The fused integrate kernel.
//
// LDS layout per cell (one wavefront):
//   lu   [SPAN + 1]  W / LU factors, row displacement; slot SPAN absorbs padding
//   ylds [NY]        stage state gathered by the flux lanes (+ fixed species, 1.0)
//   scr  [NSCR]      fluxes (F) or Jacobian partials (W); last used slot = 0.0
//   facc [NV + 1]    with the end of scr: F chunk partials (wave_eval_f)
// ---------------------------------------------------------------------------
// MODE 0: the plain integrator.  MODE 1: the same, recording the state and
// step size of every accepted step (adj->ckpt, adj->hs, adj->nsteps).  MODE 2:
// replay of recorded step sizes without error control (the frozen-step map
// the adjoint differentiates).  In modes 1 and 2, adj->k (if set) replaces the
// rate laws.  MODE 3: integrate_with_reduction
// (src/mkpp/templates/solver_variants/integrate_with_reduction.j2:51-306): F1 is
// evaluated first and sets the active rows (|F1| / (atol[t] + rtol[t] |y|) >=
// a.importance_threshold, t the row's position in the reference permutation,
// as the template indexes the tolerances); after W assembly a pass over the W
// locations turns frozen rows and columns into the identity's; stage
// right-hand sides of frozen rows are zero, which with that W makes the plain
// solves return the template's guarded zeros; dt starts at min(dt, 1 s) and is
// multiplied by the factor after every attempt.  MODE 4: the tangent-linear
// model along recorded steps: the replay of MODE 2 carries a tangent dx
// (adj->dy in, out) for a direction (dy0, adj->dk) of the initial state and
// the rate constants; per stage, W dK_s = dF_s + sum_j C_sj/h dK_j + (dJ) K_s
// with dF_s = J(Y_s) dY_s + F_k(Y_s) dk at the stages that evaluate F, and
// (dJ) K_s the derivative of J(y_n) K_s along (dx, dk); then
// dx += sum_s M_s dK_s.  The mode branches are compile-time; MODE 0 carries
// none of them.
template <class W, class Tab, int MODE = 0>
__device__ __forceinline__ void wave_integrate_cell(const DeviceArgs& a, long cell, double* lds, int lane,
                                                    const AdjArgs* adj = nullptr) {
  using T_ = WaveTables<W>;
  using C_ = typename T_::Code;
  constexpr int NV = W::NV, NB = W::NB, NG = W::NG, NGP = (NG + 63) / 64, S = Tab::S;
  double* lu = lds;
  double* ylds = lu + (W::SPAN + 1);
  double* scr = ylds + W::NY;
  double* facc = scr + W::NSCR;
  // MODE 4 only (wave_tlm_lds_doubles): y_n, dy_n, dY_s and K_s by y slot
  double* xl = facc + (NV + 1);
  double* dxl = xl + W::NY;
  double* dyl = dxl + W::NY;
  double* kvl = dyl + W::NY;
  (void)xl; (void)dxl; (void)dyl; (void)kvl;

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

  // ---------------- load ----------------
  double x[NB];
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    const int ph = valid[b] ? (int)T_::VARPHYS[rowv[b]] : 0;
    x[b] = valid[b] ? a.conc[cell + a.stride * ph] : 0.0;
  }
  if (lane < W::NF) ylds[NV + lane] = a.conc[cell + a.stride * (long)T_::FIXPHYS[lane]];
  if (lane == 0) ylds[W::ONE] = 1.0;
  const double Tk = a.temp ? a.temp[cell] : a.temp_default;
  const double rho = a.air_density ? a.air_density[cell] : a.rho_default;
  const double Mtb = (W::THIRD_BODY >= 0) ? a.conc[cell + a.stride * (long)(W::THIRD_BODY >= 0 ? W::THIRD_BODY : 0)] : rho;

  // non-finite pre-check (dispatcher.hpp:289-307) when a status array is given
  if (a.status != nullptr) {
    bool ok = true;
#pragma unroll
    for (int b = 0; b < NB; ++b) ok = ok && isfinite(x[b]);
    if (lane < W::NF) ok = ok && isfinite(ylds[NV + lane]);
    if (__ballot(!ok) != 0ull) {
      if (lane == 0) a.status[cell] = ERR_NON_FINITE_STATE;
      return;
    }
  }

  // ---------------- rate constants (once per cell and call) ----------------
  double kreg[NGP];
#pragma unroll
  for (int p = 0; p < NGP; ++p) {
    const int g = p * 64 + lane;
    kreg[p] = 0.0;
    if (g < NG) {
      const int law = T_::LAW[g];
      const int ph = T_::PHOTO[g];
      const double jv = (law == 0 && a.jvals != nullptr) ? a.jvals[cell * a.jcstride + a.jstride * ph] : 0.0;
      kreg[p] = wave_rate(law, &T_::PARAMS[g * 7], Tk, Mtb, jv);
    }
  }
  if constexpr (MODE == 1 || MODE == 2 || MODE == 4) {
    if (adj->k != nullptr) {
#pragma unroll
      for (int p = 0; p < NGP; ++p)
        if (p * 64 + lane < NG) kreg[p] = adj->k[cell + adj->kstride * (long)(p * 64 + lane)];
    }
  }
  // MODE 4: the direction of the rate constants, the tangent of the state, and
  // the constant parts of the by-slot vectors (fixed species, ONE)
  double dkreg[NGP], dx[NB];
#pragma unroll
  for (int p = 0; p < NGP; ++p) dkreg[p] = 0.0;
#pragma unroll
  for (int b = 0; b < NB; ++b) dx[b] = 0.0;
  if constexpr (MODE == 4) {
    if (adj->dk != nullptr) {
#pragma unroll
      for (int p = 0; p < NGP; ++p)
        if (p * 64 + lane < NG) dkreg[p] = adj->dk[cell + adj->kstride * (long)(p * 64 + lane)];
    }
#pragma unroll
    for (int b = 0; b < NB; ++b) {
      const int ph = valid[b] ? (int)T_::VARPHYS[rowv[b]] : 0;
      dx[b] = valid[b] ? adj->dy[cell + a.stride * ph] : 0.0;
    }
    for (int i = NV + lane; i < W::NY; i += 64) {
      const int q = i - NV;  // fixed species q, or ONE (q == NF)
      const double v = q < W::NF ? a.conc[cell + a.stride * (long)T_::FIXPHYS[q < W::NF ? q : 0]] : 1.0;
      const double d = q < W::NF ? adj->dy[cell + a.stride * (long)T_::FIXPHYS[q < W::NF ? q : 0]] : 0.0;
      xl[i] = v;
      dxl[i] = d;
      dyl[i] = d;
      kvl[i] = 0.0;
    }
  }

  auto eval_rates = [&](double (&f)[NB]) { wave_eval_f<W>(ylds, scr, kreg, valid, f, lane); };

  // W = I/(g h) - J(y) into the LU storage
  auto assemble_w = [&](double inv_g_dt) {
#pragma unroll
    for (int p = 0; p < NGP; ++p) {
      const int g = p * 64 + lane;
      if (g < NG) {
        // all three reactant entries, absent ones with power 0 (ipow = 1.0):
        // the products equal the oracle's term by term
        const int np = T_::GNPOW[g];
        const uint8_t* s = &T_::GPOWS[g * 3];
        const uint8_t* q = &T_::GPOWP[g * 3];
        const int off = T_::POFF[g];
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

  auto store_y = [&](const double (&y)[NB]) {
#pragma unroll
    for (int b = 0; b < NB; ++b)
      if (valid[b]) ylds[rowv[b]] = y[b];
  };

  // ---------------- integrate `steps` times over dt ----------------
  int status = SUCCESS;
  int acc_steps = 0, att_steps = 0;
  for (int call = 0; call < a.steps && status == SUCCESS; ++call) {
    const double dt_total = a.dt;
    const double h_min = dt_total * 1.0e-15;
    double t = 0.0, dt = MODE == 3 ? fmin(dt_total, 1.0) : dt_total * a.seed;
    bool reject_last = false, reject_more = false;
    long attempts = 0;
    int nrep = 0;
    if constexpr (MODE == 2 || MODE == 4) {
      // a recorded count outside [0, maxsteps] cannot index the cell's slice
      nrep = adj->nsteps[cell];
      if (nrep < 0 || nrep > adj->maxsteps) {
        status = ERR_CHECKPOINT_OVERFLOW;
        nrep = 0;
      }
    }
    while ((MODE == 2 || MODE == 4) ? acc_steps < nrep : t < dt_total) {
      if (a.max_attempts > 0 && attempts >= a.max_attempts) {
        status = ERR_CONVERGENCE_FAILURE;
        break;
      }
      ++attempts;
      ++att_steps;
      if constexpr (MODE == 2 || MODE == 4)
        dt = adj->hs[cell * (long)adj->maxsteps + acc_steps];
      else
        dt = fmin(dt, dt_total - t);
      const double inv_g_dt = 1.0 / (Tab::G * dt);
      // An opaque copy of the lane id: the generated code tests hundreds of
      // constant lane masks, and with a visible loop-invariant lane id the
      // compiler hoists every resulting exec mask out of the step loop and
      // spills them.  Recomputing a mask next to its use is cheaper.
      int ln = lane;
      asm volatile("" : "+v"(ln));
      double Sx[NB], K[S][NB], F[NB], rdiag[NB];
      double dK[S][NB], dF[NB];  // MODE 4 only
      bool act[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        Sx[b] = x[b];
        act[b] = valid[b];
      }
      store_y(Sx);
      if constexpr (MODE == 4) {
#pragma unroll
        for (int b = 0; b < NB; ++b)
          if (valid[b]) {
            xl[rowv[b]] = x[b];
            dxl[rowv[b]] = dx[b];
          }
      }
      // W, its LU and 1/diag.  MODE 3 runs this inside the stage loop once F1
      // (the stage-1 F, evaluated at the loop's one F site) has set the active
      // rows: a second inlined F evaluation ahead of it costs registers.
      auto factor_w = [&]() {
        assemble_w(inv_g_dt);
        if constexpr (MODE == 3) {
          uint64_t amask[NB];  // active rows, one lane mask per bank (uniform)
#pragma unroll
          for (int b = 0; b < NB; ++b) amask[b] = __ballot(act[b]);
          auto active_slot = [&](int sl) {
            uint64_t m = amask[0];
#pragma unroll
            for (int b = 1; b < NB; ++b) m = (sl >> 6) == b ? amask[b] : m;
            return ((m >> (sl & 63)) & 1ull) != 0ull;
          };
#pragma unroll
          for (int p = 0; p < W::NMWP; ++p) {
            const int e = p * 64 + ln;  // opaque lane id: keeps the table loads next to their use
            const int ad = T_::MW_ADDR[e], r = T_::MW_RSLOT[e], c = T_::MW_CSLOT[e];
            const bool ar = active_slot(r), ac = active_slot(c);
            const double v = lu[ad];
            lu[ad] = (r == c) ? (ar ? v : 1.0) : ((ar && ac) ? v : 0.0);
          }
        }
        C_::factor(lu, rb, ln);
#pragma unroll
        for (int b = 0; b < NB; ++b) rdiag[b] = valid[b] ? 1.0 / lu[rb[b] + rowv[b]] : 0.0;
      };
      if constexpr (MODE != 3) factor_w();
      // Stage loop kept rolled so that one copy of the F and solve code serves
      // every stage.  Coefficients are read by the (uniform) stage index; a
      // zero coefficient skips its term and a unit one adds K bare, exactly
      // the textual form of the stage sums.
#pragma unroll 1
      for (int s = 0; s < S; ++s) {
        double wv[NB];
        const int base = s * (s - 1) / 2;
        if (s == 0 || Tab::NEWF[s]) {
          if (s > 0) {
            double Y[NB];
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              double v = Sx[b];
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
          if constexpr (MODE == 4) {
            // dY_s = dx + sum_j A_sj dK_j by row, then dF_s = J(Y_s) dY_s + F_k dk
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              double v = dx[b];
#pragma unroll
              for (int j = 0; j < S - 1; ++j) {
                const double c = j < s ? Tab::A[base + j] : 0.0;
                if (c == 0.0) continue;
                v = (c == 1.0) ? v + dK[j][b] : v + c * dK[j][b];
              }
              if (valid[b]) dyl[rowv[b]] = v;
            }
            wave_tlm_group_jvp<W>(ylds, dyl, scr, kreg, dkreg, lane);
            wave_row_sums<W>(scr, valid, dF, lane);
          }
        }
        if constexpr (MODE == 3) {
          if (s == 0) {
            // F1 = F(S) sets the active rows (integrate_with_reduction.j2:83-90)
#pragma unroll
            for (int b = 0; b < NB; ++b) {
              const int ti = valid[b] ? (int)T_::TOLIDX[rowv[b]] : 0;
              act[b] = valid[b] && fabs(F[b]) / (a.atol[ti] + a.rtol[ti] * fabs(Sx[b])) >= a.importance_threshold;
            }
            factor_w();
          }
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
          wv[b] = act[b] ? v : 0.0;
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
        if constexpr (MODE == 4) {
          // (dJ) K_s at y_n, then W dK_s = dF_s + sum_j C_sj/h dK_j + (dJ) K_s
#pragma unroll
          for (int b = 0; b < NB; ++b)
            if (valid[b]) kvl[rowv[b]] = xs[b];
          double djk[NB], dwv[NB], dxs[NB];
          wave_tlm_group_djk<W>(xl, dxl, kvl, scr, kreg, dkreg, lane);
          wave_row_sums<W>(scr, valid, djk, lane);
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            double v = dF[b];
#pragma unroll
            for (int j = 0; j < S - 1; ++j) {
              const double c = j < s ? Tab::C[base + j] : 0.0;
              if (c == 0.0) continue;
              v = v + (c / dt) * dK[j][b];
            }
            dwv[b] = valid[b] ? v + djk[b] : 0.0;
          }
          int ld = ln;
          asm volatile("" : "+v"(ld));
          C_::forward(lu, rb, dwv, ld);
          C_::backward(lu, rb, dwv, rdiag, dxs, ld);
#pragma unroll
          for (int j = 0; j < S; ++j)
            if (j == s) {
#pragma unroll
              for (int b = 0; b < NB; ++b) dK[j][b] = dxs[b];
            }
        }
      }
      // error norm over all rows (fixed species add zeros)
      double term[NB];
#pragma unroll
      for (int b = 0; b < NB; ++b) {
        double ynew = Sx[b], yerr = 0.0;
        bool first = true;
#pragma unroll
        for (int j = 0; j < S; ++j) {
          const double m = Tab::M[j];
          if (m == 0.0) continue;
          ynew = ynew + ((m == 1.0) ? K[j][b] : m * K[j][b]);
        }
#pragma unroll
        for (int j = 0; j < S; ++j) {
          const double e = Tab::E[j];
          if (e == 0.0) continue;
          const double tt = (e == 1.0) ? K[j][b] : e * K[j][b];
          yerr = first ? tt : yerr + tt;
          first = false;
        }
        // tolerances are read where used (a few cache lines shared by every
        // cell) rather than held in registers through the step
        const int ph = valid[b] ? (int)(MODE == 3 ? T_::TOLIDX[rowv[b]] : T_::VARPHYS[rowv[b]]) : 0;
        const double sci = a.atol[ph] + a.rtol[ph] * fmax(fabs(x[b]), fabs(ynew));
        term[b] = valid[b] ? (yerr / sci) * (yerr / sci) : 0.0;
      }
      double err_sq = 0.0;
      if constexpr (MODE != 2 && MODE != 4) {
        // every row's term summed over the wavefront: each lane adds its
        // banks, then a butterfly over the 64 lanes; partners add the same
        // two values, so every lane ends with the same sum
        double sum = term[0];
#pragma unroll
        for (int b = 1; b < NB; ++b) sum += term[b];
#pragma unroll
        for (int off = 32; off >= 1; off >>= 1) sum += __shfl_xor(sum, off, 64);
        err_sq = sum;
      }
      double err = sqrt(err_sq / W::N);
      err = fmax(err, 1.0e-10);
      const double factor_ = step_factor<Tab>(err);
      if (MODE == 2 || MODE == 4 || err <= 1.0) {
        if constexpr (MODE == 1) {
          if (acc_steps >= adj->maxsteps) {
            status = ERR_CHECKPOINT_OVERFLOW;
            break;
          }
          double* ck = adj->ckpt + (cell * (long)adj->maxsteps + acc_steps) * NV;
#pragma unroll
          for (int b = 0; b < NB; ++b)
            if (valid[b]) ck[rowv[b]] = x[b];
          if (lane == 0) adj->hs[cell * (long)adj->maxsteps + acc_steps] = dt;
        }
#pragma unroll
        for (int b = 0; b < NB; ++b) {
          double msum = 0.0;
          bool first = true;
#pragma unroll
          for (int j = 0; j < S; ++j) {
            const double m = Tab::M[j];
            if (m == 0.0) continue;
            const double tt = (m == 1.0) ? K[j][b] : m * K[j][b];
            msum = first ? tt : msum + tt;
            first = false;
          }
          if (MODE != 3 || act[b]) x[b] += msum;
        }
        if constexpr (MODE == 4) {
#pragma unroll
          for (int b = 0; b < NB; ++b) {
            double msum = 0.0;
            bool first = true;
#pragma unroll
            for (int j = 0; j < S; ++j) {
              const double m = Tab::M[j];
              if (m == 0.0) continue;
              const double tt = (m == 1.0) ? dK[j][b] : m * dK[j][b];
              msum = first ? tt : msum + tt;
              first = false;
            }
            dx[b] += msum;
          }
        }
        t += dt;
        ++acc_steps;
        if constexpr (MODE == 3) {
          dt *= factor_;
        } else {
          double next_dt = fmax(h_min, fmin(dt * factor_, dt_total));
          if (reject_last) next_dt = fmin(next_dt, dt);
          reject_last = false;
          reject_more = false;
          dt = next_dt;
        }
      } else if constexpr (MODE == 3) {
        dt *= factor_;
      } else {
        dt *= reject_more ? 0.1 : factor_;
        reject_more = reject_last;
        reject_last = true;
      }
    }
  }
  // ---------------- store ----------------
  bool fin = true;
#pragma unroll
  for (int b = 0; b < NB; ++b) {
    if (valid[b]) a.conc[cell + a.stride * (long)T_::VARPHYS[rowv[b]]] = x[b];
    if constexpr (MODE == 4)
      if (valid[b]) adj->dy[cell + a.stride * (long)T_::VARPHYS[rowv[b]]] = dx[b];
    fin = fin && (!valid[b] || isfinite(x[b]));
  }
  const bool all_fin = __ballot(!fin) == 0ull;
  if (lane == 0) {
    if (status == SUCCESS && !all_fin) status = ERR_NUMERICAL_FAILURE;
    if (a.status) a.status[cell] = status;
    if (a.accepted) a.accepted[cell] += acc_steps;
    if (a.attempted) a.attempted[cell] += att_steps;
    if constexpr (MODE == 1) adj->nsteps[cell] = acc_steps;
  }
}

// XCD-aware cell order: workgroups go round-robin to the 8 XCDs, so block b
// runs on XCD b % 8.  Giving each XCD a contiguous range of cells lets the 16
// cells that share a 128-byte line of every species row (conc is
// species-major) be served by one L2 instead of eight.  gridDim.x is a multiple
// of 8, which makes the map a bijection on [0, gridDim.x).
__device__ inline long wave_cell_of_block(long base) {
  return base + (long)(blockIdx.x % 8u) * (gridDim.x / 8u) + blockIdx.x / 8u;
}

// One cell per wavefront, no loop: a launch covers cells [base, base + grid).
// Without an enclosing loop the compiler has nothing to hoist the math
// library's polynomial constants out of, so they are not kept live (and
// spilled to scratch, 8.7 KB per wave) across the whole integration.
template <class W, class Tab>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W::WAVES_PER_EU, W::WAVES_PER_EU)))
wave_integrate_kernel(DeviceArgs a, long base) {
  extern __shared__ double lds[];
  const long cell = wave_cell_of_block(base);
  if (cell < a.ncells) wave_integrate_cell<W, Tab>(a, cell, lds, threadIdx.x);
}

template <class W, class Tab>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W::WAVES_PER_EU, W::WAVES_PER_EU)))
wave_reduced_kernel(DeviceArgs a, long base) {
  extern __shared__ double lds[];
  const long cell = wave_cell_of_block(base);
  if (cell < a.ncells) wave_integrate_cell<W, Tab, 3>(a, cell, lds, threadIdx.x);
}

// Launch geometry of the one-cell-per-wave kernels: chunks of at most 2^30
// cells, each a grid that is a multiple of 8.
constexpr long WAVE_CHUNK_CELLS = 1L << 30;
inline unsigned wave_chunk_grid(long ncells, long base) {
  const long n = ncells - base < WAVE_CHUNK_CELLS ? ncells - base : WAVE_CHUNK_CELLS;
  return (unsigned)((n + 7) / 8 * 8);
}

// ---------------------------------------------------------------------------
// Evaluation kernels (one thread per cell; not on the hot path).
// ---------------------------------------------------------------------------
template <class W>
__global__ void __launch_bounds__(64) wave_eval_kernel(EvalArgs a, int what) {
  using T_ = WaveTables<W>;
  const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= a.ncells) return;
  const double Tk = a.temp ? a.temp[cell] : a.temp_default;
  const double rho = a.air_density ? a.air_density[cell] : a.rho_default;
  const double Mtb = (W::THIRD_BODY >= 0) ? a.conc[cell + a.stride * (long)(W::THIRD_BODY >= 0 ? W::THIRD_BODY : 0)] : rho;
  auto y = [&](int phys) { return T_::UNIT[phys] ? 1.0 : a.conc[cell + a.stride * phys]; };
  if (what == 0) {
    for (int i = 0; i < W::N; ++i) a.out[cell + a.out_stride * i] = 0.0;
  } else {
    for (int e = 0; e < W::N * W::N; ++e) a.out[cell + a.out_stride * (long)e] = 0.0;
  }
  for (int g = 0; g < W::NG; ++g) {
    const int law = T_::LAW[g];
    const double jv = (law == 0 && a.jvals != nullptr) ? a.jvals[cell * a.jcstride + a.jstride * T_::PHOTO[g]] : 0.0;
    const double* P = &T_::PARAMS[g * 7];
    const double k = wave_rate(law, P, Tk, Mtb, jv);
    const int np = T_::GNPOW_FULL[g];
    const uint8_t* s = &T_::GPOWS_FULL[g * 3];
    const uint8_t* q = &T_::GPOWP[g * 3];
    const int n0 = T_::GNUOFS[g], n1 = T_::GNUOFS[g + 1];
    if (what == 0) {
      double mono = 1.0;
      for (int e = 0; e < np; ++e) {
        double v = y(s[e]);
        for (int r = 1; r < q[e]; ++r) v = v * y(s[e]);
        mono = mono * v;
      }
      const double flux = k * mono;
      for (int t = n0; t < n1; ++t) {
        const long o = cell + a.out_stride * (long)T_::GNUS[t];
        a.out[o] = a.out[o] + T_::COEF[T_::GNUC[t]] * flux;
      }
    } else {
      // dk/d[M] for the third-body laws
      double dk = 0.0;
      if (law == 2) {
        const double k0 = P[0] * exp(P[2] / Tk) * pow(Tk / 300.0, P[1]);
        const double ki = P[3] * exp(P[5] / Tk) * pow(Tk / 300.0, P[4]);
        const double K0 = k0 * Mtb, r = K0 / ki, lg = log10(r), ex = 1.0 / (1.0 + lg * lg), Fb = pow(P[6], ex);
        const double dr = k0 / ki, dex = -ex * ex * 2.0 * lg / (r * log(10.0)) * dr, dF = Fb * log(P[6]) * dex;
        dk = (k0 / (1.0 + r) - K0 * dr / ((1.0 + r) * (1.0 + r))) * Fb + K0 / (1.0 + r) * dF;
      } else if (law == 3) {
        const double K2 = P[2] * exp(P[3] / Tk), k3 = P[4] * exp(P[5] / Tk), K3 = k3 * Mtb;
        dk = k3 / ((1.0 + K3 / K2) * (1.0 + K3 / K2));
      } else if (law == 4) {
        dk = P[2] * exp(P[3] / Tk);
      }
      const int d0 = T_::GDEPOFS[g], d1 = T_::GDEPOFS[g + 1];
      for (int d = d0; d < d1; ++d) {
        const int j = T_::GDEP[d];
        double part = 0.0;
        for (int e = 0; e < np; ++e) {
          if (s[e] != j) continue;
          double term = k * (double)q[e];
          double ip = 1.0;
          for (int r = 1; r < q[e]; ++r) ip = (r == 1) ? y(j) : ip * y(j);
          term = term * ip;
          for (int o = 0; o < np; ++o) {
            if (o == e) continue;
            double v = y(s[o]);
            for (int r = 1; r < q[o]; ++r) v = v * y(s[o]);
            term = term * v;
          }
          part = part + term;
        }
        if (law >= 2 && law <= 4 && j == W::THIRD_BODY) {
          double term = dk;
          for (int o = 0; o < np; ++o) {
            double v = y(s[o]);
            for (int r = 1; r < q[o]; ++r) v = v * y(s[o]);
            term = term * v;
          }
          part = part + term;
        }
        for (int t = n0; t < n1; ++t) {
          const long o = cell + a.out_stride * (long)(T_::GNUS[t] + W::N * j);
          a.out[o] = a.out[o] + T_::COEF[T_::GNUC[t]] * part;
        }
      }
    }
  }
}

// A launch above 64 KB of dynamic LDS (ts1: 66 KB per wavefront) declares its
// size to the runtime first; gfx950 gives a workgroup up to 160 KB.
template <class K>
inline void wave_allow_lds(K* kernel, size_t bytes) {
  if (bytes > 65536)
    (void)hipFuncSetAttribute(reinterpret_cast<const void*>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize,
                              (int)bytes);
}

template <class W>
int launch_wave_integrate(const DeviceArgs& a, hipStream_t s) {
  if (a.ncells <= 0) return 0;
  const size_t lds = sizeof(double) * W::LDS_DOUBLES;
  if (a.method < ROS2 || a.method > RODAS4) return -1;
  for (long base = 0; base < a.ncells; base += WAVE_CHUNK_CELLS) {
    const unsigned grid = wave_chunk_grid(a.ncells, base);
    switch (a.method) {
      case ROS2:
        wave_allow_lds(&wave_integrate_kernel<W, TabRos2>, lds);
        wave_integrate_kernel<W, TabRos2><<<grid, 64, lds, s>>>(a, base);
        break;
      case ROS3:
        wave_allow_lds(&wave_integrate_kernel<W, TabRos3>, lds);
        wave_integrate_kernel<W, TabRos3><<<grid, 64, lds, s>>>(a, base);
        break;
      case ROS4:
        wave_allow_lds(&wave_integrate_kernel<W, TabRos4>, lds);
        wave_integrate_kernel<W, TabRos4><<<grid, 64, lds, s>>>(a, base);
        break;
      case RODAS3:
        wave_allow_lds(&wave_integrate_kernel<W, TabRodas3>, lds);
        wave_integrate_kernel<W, TabRodas3><<<grid, 64, lds, s>>>(a, base);
        break;
      case RODAS4:
        wave_allow_lds(&wave_integrate_kernel<W, TabRodas4>, lds);
        wave_integrate_kernel<W, TabRodas4><<<grid, 64, lds, s>>>(a, base);
        break;
    }
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <class W>
int launch_wave_reduced(const DeviceArgs& a, hipStream_t s) {
  if (a.ncells <= 0) return 0;
  const size_t lds = sizeof(double) * W::LDS_DOUBLES;
  if (a.method < ROS2 || a.method > RODAS4) return -1;
  for (long base = 0; base < a.ncells; base += WAVE_CHUNK_CELLS) {
    const unsigned grid = wave_chunk_grid(a.ncells, base);
    switch (a.method) {
      case ROS2:
        wave_allow_lds(&wave_reduced_kernel<W, TabRos2>, lds);
        wave_reduced_kernel<W, TabRos2><<<grid, 64, lds, s>>>(a, base);
        break;
      case ROS3:
        wave_allow_lds(&wave_reduced_kernel<W, TabRos3>, lds);
        wave_reduced_kernel<W, TabRos3><<<grid, 64, lds, s>>>(a, base);
        break;
      case ROS4:
        wave_allow_lds(&wave_reduced_kernel<W, TabRos4>, lds);
        wave_reduced_kernel<W, TabRos4><<<grid, 64, lds, s>>>(a, base);
        break;
      case RODAS3:
        wave_allow_lds(&wave_reduced_kernel<W, TabRodas3>, lds);
        wave_reduced_kernel<W, TabRodas3><<<grid, 64, lds, s>>>(a, base);
        break;
      case RODAS4:
        wave_allow_lds(&wave_reduced_kernel<W, TabRodas4>, lds);
        wave_reduced_kernel<W, TabRodas4><<<grid, 64, lds, s>>>(a, base);
        break;
    }
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <class W>
int launch_wave_eval(const EvalArgs& a, hipStream_t s, int what) {
  if (a.ncells <= 0) return 0;
  wave_eval_kernel<W><<<dim3((unsigned)((a.ncells + 63) / 64)), dim3(64), 0, s>>>(a, what);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace mkpp

// The table bundle of a generated wave mechanism (any unit that runs its kernel
// needs it: the mechanism's own unit and the batch unit).
#define MKPP_WAVE_TABLES(NAME)                                                                              \
  namespace mkpp {                                                                                          \
  template <>                                                                                               \
  struct WaveTables<gen::NAME##_wave::Wave> {                                                               \
    static constexpr const int16_t* R = gen::NAME##_wave::R;                                               \
    static constexpr const uint8_t* SLOT = gen::NAME##_wave::SLOT;                                          \
    static constexpr const uint8_t* ROWOF = gen::NAME##_wave::ROWOF;                                        \
    static constexpr const uint8_t* LCNT = gen::NAME##_wave::LCNT;                                          \
    static constexpr const uint16_t* LOFS = gen::NAME##_wave::LOFS;                                         \
    static constexpr const uint8_t* LROW = gen::NAME##_wave::LROW;                                          \
    static constexpr const uint16_t* UOFS = gen::NAME##_wave::UOFS;                                         \
    static constexpr const uint8_t* UCOL = gen::NAME##_wave::UCOL;                                          \
    static constexpr const uint64_t* FMASK = gen::NAME##_wave::FMASK;                                       \
    static constexpr const uint64_t* BMASK = gen::NAME##_wave::BMASK;                                       \
    static constexpr const uint8_t* LAW = gen::NAME##_wave::LAW;                                            \
    static constexpr const int16_t* PHOTO = gen::NAME##_wave::PHOTO;                                        \
    static constexpr const double* PARAMS = gen::NAME##_wave::PARAMS;                                       \
    static constexpr const uint8_t* GFAC = gen::NAME##_wave::GFAC;                                          \
    static constexpr const uint8_t* GNPOW = gen::NAME##_wave::GNPOW;                                        \
    static constexpr const uint8_t* GPOWS = gen::NAME##_wave::GPOWS;                                        \
    static constexpr const uint8_t* GPOWK = gen::NAME##_wave::GPOWK;                                        \
    static constexpr const uint8_t* GPOWP = gen::NAME##_wave::GPOWP;                                        \
    static constexpr const uint16_t* POFF = gen::NAME##_wave::POFF;                                         \
    static constexpr const double* COEF = gen::NAME##_wave::COEF;                                           \
    static constexpr const uint16_t* FOFS = gen::NAME##_wave::FOFS;                                         \
    static constexpr const uint16_t* FG = gen::NAME##_wave::FG;                                             \
    static constexpr const uint8_t* FC = gen::NAME##_wave::FC;                                              \
    static constexpr const uint16_t* WADDR = gen::NAME##_wave::WADDR;                                       \
    static constexpr const uint8_t* WDIAG = gen::NAME##_wave::WDIAG;                                        \
    static constexpr const uint16_t* WOFS = gen::NAME##_wave::WOFS;                                         \
    static constexpr const uint16_t* WPI = gen::NAME##_wave::WPI;                                           \
    static constexpr const uint8_t* WCI = gen::NAME##_wave::WCI;                                            \
    static constexpr const uint8_t* VARPHYS = gen::NAME##_wave::VARPHYS;                                    \
    static constexpr const uint8_t* FIXPHYS = gen::NAME##_wave::FIXPHYS;                                    \
    static constexpr const uint8_t* UNIT = gen::NAME##_wave::UNIT;                                          \
    static constexpr const uint8_t* GNPOW_FULL = gen::NAME##_wave::GNPOW;                                   \
    static constexpr const uint8_t* GPOWS_FULL = gen::NAME##_wave::GPOWS_PHYS;                              \
    static constexpr const uint16_t* GNUOFS = gen::NAME##_wave::GNUOFS;                                     \
    static constexpr const uint8_t* GNUS = gen::NAME##_wave::GNUS;                                          \
    static constexpr const uint8_t* GNUC = gen::NAME##_wave::GNUC;                                          \
    static constexpr const uint16_t* GDEPOFS = gen::NAME##_wave::GDEPOFS;                                   \
    static constexpr const uint8_t* GDEP = gen::NAME##_wave::GDEP;                                          \
    static constexpr const uint32_t* CF_TERM = gen::NAME##_wave::CF_TERM;                                   \
    static constexpr const uint16_t* CF_PART = gen::NAME##_wave::CF_PART;                                    \
    static constexpr const uint32_t* CW_TERM = gen::NAME##_wave::CW_TERM;                                   \
    static constexpr const uint16_t* CW_ADDR = gen::NAME##_wave::CW_ADDR;                                   \
    static constexpr const uint16_t* GNVOFS = gen::NAME##_wave::GNVOFS;                                     \
    static constexpr const uint8_t* GNVR = gen::NAME##_wave::GNVR;                                          \
    static constexpr const uint8_t* GNVC = gen::NAME##_wave::GNVC;                                          \
    static constexpr const uint16_t* SPOFS = gen::NAME##_wave::SPOFS;                                       \
    static constexpr const uint16_t* SPI = gen::NAME##_wave::SPI;                                           \
    static constexpr const uint32_t* CS_TERM = gen::NAME##_wave::CS_TERM;                                   \
    static constexpr const uint16_t* CS_PART = gen::NAME##_wave::CS_PART;                                   \
    static constexpr const uint16_t* CS_FIX = gen::NAME##_wave::CS_FIX;                                     \
    static constexpr const uint16_t* CG_TERM = gen::NAME##_wave::CG_TERM;                                   \
    static constexpr const uint16_t* CG_GRP = gen::NAME##_wave::CG_GRP;                                     \
    static constexpr const uint16_t* MW_ADDR = gen::NAME##_wave::MW_ADDR;                                   \
    static constexpr const uint8_t* MW_RSLOT = gen::NAME##_wave::MW_RSLOT;                                  \
    static constexpr const uint8_t* MW_CSLOT = gen::NAME##_wave::MW_CSLOT;                                  \
    static constexpr const uint8_t* TOLIDX = gen::NAME##_wave::TOLIDX;                                      \
    using Code = gen::NAME##_wave::Code;                                                                    \
  };                                                                                                        \
  }

// Registers a generated wave mechanism without the differentiation kernels
// (the diff entry points return -103 for it).
#define MKPP_WAVE_MECHANISM_NODIFF(NAME, F0, F1, F2, F3, F4)                                                  \
  MKPP_WAVE_TABLES(NAME)                                                                                    \
  namespace mkpp {                                                                                          \
  static int wave_eval_rates_##NAME(const EvalArgs& a, hipStream_t s) {                                     \
    return launch_wave_eval<gen::NAME##_wave::Wave>(a, s, 0);                                               \
  }                                                                                                         \
  static int wave_eval_jac_##NAME(const EvalArgs& a, hipStream_t s) {                                       \
    return launch_wave_eval<gen::NAME##_wave::Wave>(a, s, 1);                                               \
  }                                                                                                         \
  const MechOps* mkpp_ops_##NAME() {                                                                        \
    using W = gen::NAME##_wave::Wave;                                                                       \
    static const MechOps ops = {#NAME, W::N, W::NUM_REACTIONS, W::NPH, W::NPH, 0,                          \
                                W::SUPPORTS_RUNTIME_ENVIRONMENT, W::SPECIES,                                \
                                &launch_wave_integrate<W>, &wave_eval_rates_##NAME, &wave_eval_jac_##NAME, \
                                {F0, F1, F2, F3, F4}, nullptr, W::NG, W::NV,                               \
                                &launch_wave_reduced<W>};                                                   \
    return &ops;                                                                                            \
  }                                                                                                         \
  }

// Registers a generated wave mechanism: defines mkpp_ops_<name>().
#define MKPP_WAVE_MECHANISM(NAME, F0, F1, F2, F3, F4)                                                         \
  MKPP_WAVE_TABLES(NAME)                                                                                    \
  namespace mkpp {                                                                                          \
  static int wave_eval_rates_##NAME(const EvalArgs& a, hipStream_t s) {                                     \
    return launch_wave_eval<gen::NAME##_wave::Wave>(a, s, 0);                                               \
  }                                                                                                         \
  static int wave_eval_jac_##NAME(const EvalArgs& a, hipStream_t s) {                                       \
    return launch_wave_eval<gen::NAME##_wave::Wave>(a, s, 1);                                               \
  }                                                                                                         \
  const MechOps* mkpp_ops_##NAME() {                                                                        \
    using W = gen::NAME##_wave::Wave;                                                                       \
    static const MechOps ops = {#NAME, W::N, W::NUM_REACTIONS, W::NPH, W::NPH, 0,                          \
                                W::SUPPORTS_RUNTIME_ENVIRONMENT, W::SPECIES,                                \
                                &launch_wave_integrate<W>, &wave_eval_rates_##NAME, &wave_eval_jac_##NAME, \
                                {F0, F1, F2, F3, F4}, &launch_wave_diff<W>, W::NG, W::NV,                  \
                                &launch_wave_reduced<W>};                                                   \
    return &ops;                                                                                            \
  }                                                                                                         \
  }
