// Device-side pieces shared by every mechanism: Rosenbrock tableaux, the
// per-cell step controller and the register-resident integrator used by the
// small mechanisms (one cell per lane).
//
// The control flow restates the reference's generated integrate()
// (src/mkpp/templates/solver_variants/integrate.j2:52-381):
//   dt = dt_total * seed; while (t < dt_total) { dt = min(dt, dt_total - t);
//   W = I/(g dt) - J(S); LU; S stages; error norm over all N species;
//   factor = clamp(0.9 err^(-1/ELO), 0.2, 6); accept/reject with MICM's
//   no-growth-after-reject and x0.1-after-two-rejects rules }.
// One deliberate addition: a cell stops after max_attempts attempted steps
// and reports ERR_CONVERGENCE_FAILURE, so no kernel can spin forever on a
// collapsing step size (the reference loops without bound there).
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>

namespace mkpp {

// src/mkpp_host/include/mkpp_host/dispatcher.hpp:15-21
enum CellErrorCode : int {
  SUCCESS = 0,
  ERR_SINGULAR_MATRIX = 1,
  ERR_NON_FINITE_STATE = 2,
  ERR_CONVERGENCE_FAILURE = 3,
  ERR_NUMERICAL_FAILURE = 4,
  ERR_CHECKPOINT_OVERFLOW = 5,  // more accepted steps than the checkpoint buffer holds
};

enum Method : int { ROS2 = 0, ROS3 = 1, ROS4 = 2, RODAS3 = 3, RODAS4 = 4 };

// Arguments of one integrate launch.  Layout is the C-ABI's LayoutLeft:
// conc[cell + stride * species], jvals[cell * jcstride + jstride * slot]
// (jcstride 1: the C ABI's per-cell LayoutLeft forcing; 0: one vector shared
// by every cell; any other value: a strided [cells x slots] view).
struct DeviceArgs {
  double* conc;
  long ncells;
  long stride;
  const double* temp;         // [ncells] or null (temp_default)
  const double* air_density;  // [ncells] or null (rho_default)
  const double* jvals;        // rate parameters (see above) or null (zeros)
  long jstride;
  long jcstride = 1;
  double temp_default = 300.0;
  double rho_default = 2.4476e19;
  double dt;
  int steps;
  int method;
  double seed;
  const double* atol;  // [N] physical order, device memory
  const double* rtol;  // [N]
  int* status;         // [ncells] or null
  int* accepted;       // [ncells] or null: accepted steps summed over calls
  int* attempted;      // [ncells] or null
  long max_attempts;   // per integrate() call; <= 0 means unbounded
  double importance_threshold;  // reduction variant only: species below it are frozen
  const double* cloud_water = nullptr;  // [ncells] cloud liquid water, or null (clw_default)
  double clw_default = 0.0;
};

// Differentiation arguments (checkpointed forward, frozen-step replay, adjoint).
// Checkpoints are cell-major so that one wavefront writes a contiguous run:
//   ckpt[(cell * maxsteps + n) * NV + v]  state (variable species, solver row
//                                         order) at the start of accepted step n
//   hs[cell * maxsteps + n]               its step size
struct AdjArgs {
  const double* k;  // [NG x kstride] rate constants used instead of the rate laws, or null
  long kstride;
  double* ckpt;
  double* hs;
  int* nsteps;  // [ncells] accepted steps recorded
  int maxsteps;
  double* lambda;  // adjoint: [N x stride] physical order; in dL/dy(T), out dL/dy(0)
  double* kbar;    // adjoint: [NG x kstride] out dL/dk
  double* dy;         // tangent-linear: [N x stride] physical order; in dy(0), out dy(T)
  const double* dk;   // tangent-linear: [NG x kstride] direction of the rate constants, or null (0)
};

// ---------------------------------------------------------------------------
// Tableaux: Sandu et al. (1997) coefficient sets as carried by KPP and by the
// reference (src/mkpp/rosenbrock.py:27-133).  A and C are row-wise strictly
// lower triangular: A(i, j) = A[i*(i-1)/2 + j].
// ---------------------------------------------------------------------------
struct TabRos2 {
  static constexpr int S = 2;
  static constexpr double G = 1.7071067811865475244;
  static constexpr double ELO = 2.0;
  static constexpr double A[1] = {1.7071067811865475244};
  static constexpr double C[1] = {-1.1715728752538099024};
  static constexpr double M[2] = {0.87867965644035742774, 0.29289321881345247560};
  static constexpr double E[2] = {0.29289321881345247560, 0.29289321881345247560};
  static constexpr bool NEWF[2] = {true, true};
};
struct TabRos3 {
  static constexpr int S = 3;
  static constexpr double G = 0.43586652150845899942;
  static constexpr double ELO = 3.0;
  static constexpr double A[3] = {1.0, 1.0, 0.0};
  static constexpr double C[3] = {-1.0156171083877702092, 4.0759956452537699825, 9.2076794298330791242};
  static constexpr double M[3] = {1.0, 6.1697947043828245593, -0.42772256543218573326};
  static constexpr double E[3] = {0.5, -2.9079558716805469822, 0.22354069897811569627};
  static constexpr bool NEWF[3] = {true, true, false};
};
struct TabRos4 {
  static constexpr int S = 4;
  static constexpr double G = 0.5728200000000000;
  static constexpr double ELO = 4.0;
  static constexpr double A[6] = {2.0, 1.867943637803922, 0.2344449711399156, 1.867943637803922, 0.2344449711399156, 0.0};
  static constexpr double C[6] = {-7.137615036412310, 2.580708087951457, 0.6515950076447975,
                                  -2.137148994382534, -0.3214669691237626, -0.6949742501781779};
  static constexpr double M[4] = {2.255570073418735, 0.2870493262186792, 0.4353179431840180, 1.093502252409163};
  static constexpr double E[4] = {-0.2815431932141155, -0.07276199124938920, -0.1082196201495311, -1.093502252409163};
  static constexpr bool NEWF[4] = {true, true, true, false};
};
struct TabRodas3 {
  static constexpr int S = 4;
  static constexpr double G = 0.5;
  static constexpr double ELO = 3.0;
  static constexpr double A[6] = {0.0, 2.0, 0.0, 2.0, 0.0, 1.0};
  static constexpr double C[6] = {4.0, 1.0, -1.0, 1.0, -1.0, -8.0 / 3.0};
  static constexpr double M[4] = {2.0, 0.0, 1.0, 1.0};
  static constexpr double E[4] = {0.0, 0.0, 0.0, 1.0};
  static constexpr bool NEWF[4] = {true, false, true, true};
};
struct TabRodas4 {
  static constexpr int S = 6;
  static constexpr double G = 0.25;
  static constexpr double ELO = 4.0;
  static constexpr double A[15] = {1.544000000000000, 0.9466785280815826, 0.2557011698983284, 3.314825187068521,
                                   2.896124015972201, 0.9986419139977817, 1.221224509226641,  6.019134481288629,
                                   12.53708332932087, -0.6878860361058950, 1.221224509226641, 6.019134481288629,
                                   12.53708332932087, -0.6878860361058950, 1.0};
  static constexpr double C[15] = {-5.668800000000000, -2.430093356833875, -0.2063599157091915, -0.1073529058151375,
                                   -9.594562251023355, -20.47028614809616, 7.496443313967647,  -10.24680431464352,
                                   -33.99990352819905, 11.70890893206160,  8.083246795921522,  -7.981132988064893,
                                   -31.52159432874371, 16.31930543123136,  -6.058818238834054};
  static constexpr double M[6] = {1.221224509226641, 6.019134481288629, 12.53708332932087, -0.6878860361058950, 1.0, 1.0};
  static constexpr double E[6] = {0.0, 0.0, 0.0, 0.0, 0.0, 1.0};
  static constexpr bool NEWF[6] = {true, true, true, true, true, true};
};

// Y_s = S + A_s1 K_1 + ... in the template's textual form (integrate.j2:147-174):
// zero coefficients are dropped, a coefficient of exactly 1.0 is a bare term.
template <class Tab, int s>
__device__ __forceinline__ double stage_state(double S, const double (&K)[Tab::S], int /*unused*/ = 0) {
  double v = S;
#pragma unroll
  for (int j = 0; j < s; ++j) {
    constexpr int base = s * (s - 1) / 2;
    const double a = Tab::A[base + j];
    if (a == 0.0) continue;
    v = (a == 1.0) ? v + K[j] : v + a * K[j];
  }
  return v;
}

// Step-size factor of the controller (integrate.j2:321-333).
template <class Tab>
__device__ __forceinline__ double step_factor(double err) {
  double f;
  if constexpr (Tab::ELO == 2.0) {
    f = 0.9 / sqrt(err);
  } else if constexpr (Tab::ELO == 3.0) {
    f = 0.9 / cbrt(err);
  } else {
    f = 0.9 * pow(err, -(1.0 / Tab::ELO));
  }
  return fmax(0.2, fmin(f, 6.0));
}

// Register-resident integrate(dt_total) of one cell (one lane).
// x: variable species in logical order, updated in place.
template <class M, class Tab>
__device__ __forceinline__ int integrate_cell(double (&x)[M::NV], const double (&yf)[M::NF], const double* k,
                                              const double* dk, double dt_total, double seed,
                                              const double* __restrict__ atol, const double* __restrict__ rtol,
                                              long max_attempts, int& accepted, int& attempted) {
  constexpr int NV = M::NV, S = Tab::S;
  const double g = Tab::G;
  const double h_min = dt_total * 1.0e-15;
  double t = 0.0, dt = dt_total * seed;
  bool reject_last = false, reject_more = false;
  long attempts = 0;
  double Sx[NV], K[S][NV], F[NV], LU[M::NLU];
  while (t < dt_total) {
    if (max_attempts > 0 && attempts >= max_attempts) return ERR_CONVERGENCE_FAILURE;
    ++attempts;
    ++attempted;
    dt = fmin(dt, dt_total - t);
    const double inv_g_dt = 1.0 / (g * dt);
#pragma unroll
    for (int i = 0; i < NV; ++i) Sx[i] = x[i];
    M::jac_w(Sx, yf, k, dk, inv_g_dt, LU);
    M::factor(LU);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      double rhs[NV];
      if (s == 0) {
        M::rates(Sx, yf, k, F);
#pragma unroll
        for (int i = 0; i < NV; ++i) rhs[i] = F[i];
      } else {
        const int base = s * (s - 1) / 2;
        if (Tab::NEWF[s]) {
          double Y[NV];
#pragma unroll
          for (int i = 0; i < NV; ++i) {
            double v = Sx[i];
#pragma unroll
            for (int j = 0; j < s; ++j) {
              const double a = Tab::A[base + j];
              if (a == 0.0) continue;
              v = (a == 1.0) ? v + K[j][i] : v + a * K[j][i];
            }
            Y[i] = v;
          }
          M::rates(Y, yf, k, F);
        }
#pragma unroll
        for (int i = 0; i < NV; ++i) {
          double v = F[i];
#pragma unroll
          for (int j = 0; j < s; ++j) {
            const double c = Tab::C[base + j];
            if (c == 0.0) continue;
            v = v + (c / dt) * K[j][i];
          }
          rhs[i] = v;
        }
      }
      M::solve(LU, rhs, K[s]);
    }
    // error norm over all N species; fixed species add exact zeros (integrate.j2:265-322)
    double err_sq = 0.0;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double ynew = Sx[i];
      double yerr = 0.0;
      bool first = true;
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const double m = Tab::M[j];
        if (m == 0.0) continue;
        ynew = ynew + ((m == 1.0) ? K[j][i] : m * K[j][i]);
      }
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const double e = Tab::E[j];
        if (e == 0.0) continue;
        const double term = (e == 1.0) ? K[j][i] : e * K[j][i];
        yerr = first ? term : yerr + term;
        first = false;
      }
      const double ymax = fmax(fabs(x[i]), fabs(ynew));
      // per-species tolerances are read where used (uniform loads), not held in registers
      const double sci = atol[M::var_phys(i)] + rtol[M::var_phys(i)] * ymax;
      err_sq += (yerr / sci) * (yerr / sci);
    }
    double err = sqrt(err_sq / M::N);
    err = fmax(err, 1.0e-10);
    const double factor = step_factor<Tab>(err);
    if (err <= 1.0) {
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        double msum = 0.0;
        bool first = true;
#pragma unroll
        for (int j = 0; j < S; ++j) {
          const double m = Tab::M[j];
          if (m == 0.0) continue;
          const double term = (m == 1.0) ? K[j][i] : m * K[j][i];
          msum = first ? term : msum + term;
          first = false;
        }
        x[i] += msum;
      }
      t += dt;
      ++accepted;
      double next_dt = fmax(h_min, fmin(dt * factor, dt_total));
      if (reject_last) next_dt = fmin(next_dt, dt);
      reject_last = false;
      reject_more = false;
      dt = next_dt;
    } else {
      dt *= reject_more ? 0.1 : factor;
      reject_more = reject_last;
      reject_last = true;
    }
  }
  return SUCCESS;
}

// integrate_with_reduction(dt_total, importance_threshold) of one cell (one
// lane), after src/mkpp/templates/solver_variants/integrate_with_reduction.j2:51-306:
// every attempt evaluates F1 at the step's start state and keeps species i
// active when |F1_i| / (atol[t] + rtol[t] |y_i|) >= threshold, t being the
// species' position in the reference's permutation (the template indexes the
// tolerance arrays by that position, :86-88 and :237).  A frozen species' row
// and column of W become the identity's (:92-110) and its stage increments are
// zero (macros/lu_solve_forward.j2:30-31, lu_solve_backward.j2:30-31).  A
// frozen row of that W has no off-diagonal entries and a frozen column none
// either, so the plain factorization leaves them so and a plain solve with the
// frozen right-hand sides set to zero returns exactly the guarded solve's
// zeros; the LU and solve code are the plain integrator's.  Controller
// (:55, :274-305): dt starts at min(dt_total, 1 s) and is multiplied by the
// clamped factor after every attempt, accepted or not; no h_min.
template <class M, class Tab>
__device__ __forceinline__ int integrate_cell_reduced(double (&x)[M::NV], const double (&yf)[M::NF], const double* k,
                                                      const double* dk, double dt_total, double threshold,
                                                      const double* __restrict__ atol,
                                                      const double* __restrict__ rtol, long max_attempts,
                                                      int& accepted, int& attempted) {
  constexpr int NV = M::NV, S = Tab::S;
  const double g = Tab::G;
  double t = 0.0, dt = fmin(dt_total, 1.0);
  long attempts = 0;
  double Sx[NV], K[S][NV], F[NV], LU[M::NLU];
  bool act[NV];
  while (t < dt_total) {
    if (max_attempts > 0 && attempts >= max_attempts) return ERR_CONVERGENCE_FAILURE;
    ++attempts;
    ++attempted;
    dt = fmin(dt, dt_total - t);
    const double inv_g_dt = 1.0 / (g * dt);
#pragma unroll
    for (int i = 0; i < NV; ++i) Sx[i] = x[i];
    M::rates(Sx, yf, k, F);
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      const int ti = M::tol_index(i);
      act[i] = fabs(F[i]) / (atol[ti] + rtol[ti] * fabs(Sx[i])) >= threshold;
    }
    M::jac_w(Sx, yf, k, dk, inv_g_dt, LU);
#pragma unroll
    for (int e = 0; e < M::NWM; ++e) {
      const int sl = M::wm_slot(e), r = M::wm_row(e), c = M::wm_col(e);
      LU[sl] = (r == c) ? (act[r] ? LU[sl] : 1.0) : ((act[r] && act[c]) ? LU[sl] : 0.0);
    }
    M::factor(LU);
#pragma unroll
    for (int s = 0; s < S; ++s) {
      double rhs[NV];
      if (s == 0) {
#pragma unroll
        for (int i = 0; i < NV; ++i) rhs[i] = F[i];
      } else {
        const int base = s * (s - 1) / 2;
        if (Tab::NEWF[s]) {
          double Y[NV];
#pragma unroll
          for (int i = 0; i < NV; ++i) {
            double v = Sx[i];
#pragma unroll
            for (int j = 0; j < s; ++j) {
              const double a = Tab::A[base + j];
              if (a == 0.0) continue;
              v = (a == 1.0) ? v + K[j][i] : v + a * K[j][i];
            }
            Y[i] = v;
          }
          M::rates(Y, yf, k, F);
        }
#pragma unroll
        for (int i = 0; i < NV; ++i) {
          double v = F[i];
#pragma unroll
          for (int j = 0; j < s; ++j) {
            const double c = Tab::C[base + j];
            if (c == 0.0) continue;
            v = v + (c / dt) * K[j][i];
          }
          rhs[i] = v;
        }
      }
#pragma unroll
      for (int i = 0; i < NV; ++i) rhs[i] = act[i] ? rhs[i] : 0.0;
      M::solve(LU, rhs, K[s]);
    }
    double err_sq = 0.0;
#pragma unroll
    for (int i = 0; i < NV; ++i) {
      double ynew = Sx[i];
      double yerr = 0.0;
      bool first = true;
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const double m = Tab::M[j];
        if (m == 0.0) continue;
        ynew = ynew + ((m == 1.0) ? K[j][i] : m * K[j][i]);
      }
#pragma unroll
      for (int j = 0; j < S; ++j) {
        const double e = Tab::E[j];
        if (e == 0.0) continue;
        const double term = (e == 1.0) ? K[j][i] : e * K[j][i];
        yerr = first ? term : yerr + term;
        first = false;
      }
      const int ti = M::tol_index(i);
      const double sci = atol[ti] + rtol[ti] * fmax(fabs(x[i]), fabs(ynew));
      err_sq += (yerr / sci) * (yerr / sci);
    }
    double err = sqrt(err_sq / M::N);
    err = fmax(err, 1.0e-10);
    const double factor = step_factor<Tab>(err);
    if (err <= 1.0) {
#pragma unroll
      for (int i = 0; i < NV; ++i) {
        double msum = 0.0;
        bool first = true;
#pragma unroll
        for (int j = 0; j < S; ++j) {
          const double m = Tab::M[j];
          if (m == 0.0) continue;
          const double term = (m == 1.0) ? K[j][i] : m * K[j][i];
          msum = first ? term : msum + term;
          first = false;
        }
        if (act[i]) x[i] += msum;
      }
      t += dt;
      ++accepted;
    }
    dt *= factor;
  }
  return SUCCESS;
}

__device__ __forceinline__ bool all_finite(const double* v, int n) {
  bool ok = true;
  for (int i = 0; i < n; ++i) ok = ok && isfinite(v[i]);
  return ok;
}

// One lane per cell: load, hoist rate constants, integrate `steps` times, store.
// Shared by the plain launch below and the persistent batch kernel; REDUCED
// selects the reduction variant's integrator.
template <class M, class Tab, bool REDUCED = false>
__device__ __forceinline__ void integrate_lane_cell(const DeviceArgs& a, long cell) {
  double x[M::NV], yf[M::NF], jv[M::NPHOTO];
  double k[M::NK], dk[M::NK];
#pragma unroll
  for (int v = 0; v < M::NV; ++v) {
    x[v] = a.conc[cell + a.stride * M::var_phys(v)];
  }
#pragma unroll
  for (int q = 0; q < M::NF; ++q) yf[q] = (q < M::NFIXED) ? a.conc[cell + a.stride * M::fixed_phys(q)] : 0.0;
#pragma unroll
  for (int p = 0; p < M::NPHOTO; ++p)
    jv[p] = (a.jvals != nullptr && p < M::NUM_RUNTIME_PARAMETERS) ? a.jvals[cell * a.jcstride + a.jstride * p] : 0.0;
  const double T = a.temp ? a.temp[cell] : a.temp_default;
  const double rho = a.air_density ? a.air_density[cell] : a.rho_default;
  double Mtb = rho;
#pragma unroll
  for (int q = 0; q < M::NFIXED; ++q)
    if (M::fixed_phys(q) == M::THIRD_BODY) Mtb = yf[q];
  int status = SUCCESS;
  if (a.status != nullptr && !(all_finite(x, M::NV) && all_finite(yf, M::NFIXED))) {
    // batched driver's pre-check (dispatcher.hpp:289-307): flag and skip
    a.status[cell] = ERR_NON_FINITE_STATE;
    return;
  }
  M::rate_constants(T, Mtb, jv, a.cloud_water ? a.cloud_water[cell] : a.clw_default, k, dk);
  int acc = 0, att = 0;
  for (int step = 0; step < a.steps && status == SUCCESS; ++step) {
    if constexpr (REDUCED)
      status = integrate_cell_reduced<M, Tab>(x, yf, k, dk, a.dt, a.importance_threshold, a.atol, a.rtol,
                                              a.max_attempts, acc, att);
    else
      status = integrate_cell<M, Tab>(x, yf, k, dk, a.dt, a.seed, a.atol, a.rtol, a.max_attempts, acc, att);
  }
#pragma unroll
  for (int v = 0; v < M::NV; ++v) a.conc[cell + a.stride * M::var_phys(v)] = x[v];
  if (status == SUCCESS && !all_finite(x, M::NV)) status = ERR_NUMERICAL_FAILURE;
  if (a.status) a.status[cell] = status;
  if (a.accepted) a.accepted[cell] += acc;
  if (a.attempted) a.attempted[cell] += att;
}

template <class M, class Tab>
__global__ void __launch_bounds__(128) integrate_lane_kernel(DeviceArgs a) {
  const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell < a.ncells) integrate_lane_cell<M, Tab>(a, cell);
}

template <class M, class Tab>
__global__ void __launch_bounds__(128) integrate_lane_reduced_kernel(DeviceArgs a) {
  const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell < a.ncells) integrate_lane_cell<M, Tab, true>(a, cell);
}

// Evaluation kernels (mkpp_evaluate_rates / mkpp_evaluate_jacobian).
struct EvalArgs {
  const double* conc;
  long ncells;
  long stride;
  const double* temp;
  const double* air_density;
  const double* jvals;
  long jstride;
  long jcstride = 1;
  double temp_default = 300.0;
  double rho_default = 2.4476e19;
  double* out;  // rates: [ncells x N]; jacobian: [ncells x N x N], LayoutLeft
  long out_stride;
  const double* cloud_water = nullptr;  // [ncells] cloud liquid water, or null (clw_default)
  double clw_default = 0.0;
};

template <class M>
__device__ __forceinline__ void eval_setup(const EvalArgs& a, long cell, double (&y)[M::N], double* k, double* dk) {
  double jv[M::NPHOTO];
#pragma unroll
  for (int i = 0; i < M::N; ++i) y[i] = a.conc[cell + a.stride * i];
#pragma unroll
  for (int p = 0; p < M::NPHOTO; ++p)
    jv[p] = (a.jvals != nullptr && p < M::NUM_RUNTIME_PARAMETERS) ? a.jvals[cell * a.jcstride + a.jstride * p] : 0.0;
  const double T = a.temp ? a.temp[cell] : a.temp_default;
  const double rho = a.air_density ? a.air_density[cell] : a.rho_default;
  const double Mtb = M::THIRD_BODY >= 0 ? y[M::THIRD_BODY >= 0 ? M::THIRD_BODY : 0] : rho;
  M::rate_constants(T, Mtb, jv, a.cloud_water ? a.cloud_water[cell] : a.clw_default, k, dk);
}

template <class M>
__global__ void __launch_bounds__(64) eval_rates_kernel(EvalArgs a) {
  const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= a.ncells) return;
  double y[M::N], k[M::NK], dk[M::NK], f[M::N];
  eval_setup<M>(a, cell, y, k, dk);
  M::rates_full(y, k, f);
#pragma unroll
  for (int i = 0; i < M::N; ++i) a.out[cell + a.out_stride * i] = f[i];
}

template <class M>
__global__ void __launch_bounds__(64) eval_jacobian_kernel(EvalArgs a) {
  const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= a.ncells) return;
  double y[M::N], k[M::NK], dk[M::NK];
  eval_setup<M>(a, cell, y, k, dk);
  M::jac_full(y, k, dk, a.out + cell, a.out_stride);
}

}  // namespace mkpp
