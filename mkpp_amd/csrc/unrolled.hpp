// Validation backend: integrate() over dense unrolled LU and solves.
//
// The reference can build its integrator against the straight-line symbolic
// LU of device_impl/factorize.hpp.j2:13-18 and solve.hpp.j2:12-21 instead of
// the compact plan (MKPP_USE_UNROLLED_REFERENCE,
// src/mkpp/templates/solver_variants/integrate.j2:136-141, 255-259), to check
// the plan against the symbolic algebra it was derived from.  This kernel does
// the same for a lane-per-cell mechanism M: one cell per lane, every species
// in the solver permutation (fixed species included, as the reference's N x N
// matrices hold them), W assembled densely from the physical Jacobian
// (integrate.j2:103-137), the generated Unrolled::factorize / solve
// (mkpp_amd.gen.emit_unrolled), and the stage sums, error norm and step
// control of integrate_cell.  Its results on the variable species equal the
// compact-plan kernel's bit for bit: every LU entry and solve row takes the
// same operations in the same order, and the fixed species' rows and columns
// only add or subtract exact zeros.  Not tuned for speed.
#pragma once
#include <hip/hip_runtime.h>

#include "mkpp_device.hpp"

#define MKPP_UNROLLED_FN __device__ __forceinline__

namespace mkpp {

template <class M, class U, class Tab>
__device__ __forceinline__ int integrate_cell_unrolled(double (&y)[U::N], const double* k, const double* dk,
                                                       double dt_total, double seed,
                                                       const double* __restrict__ atol,
                                                       const double* __restrict__ rtol, long max_attempts,
                                                       int& accepted, int& attempted) {
  constexpr int N = U::N, S = Tab::S;
  const double g = Tab::G;
  const double h_min = dt_total * 1.0e-15;
  double t = 0.0, dt = dt_total * seed;
  bool reject_last = false, reject_more = false;
  long attempts = 0;
  double Sx[N], K[S][N], F[N], fph[N], yph[N], J[N * N], W[N * N], LU[N * N];
  while (t < dt_total) {
    if (max_attempts > 0 && attempts >= max_attempts) return ERR_CONVERGENCE_FAILURE;
    ++attempts;
    ++attempted;
    dt = fmin(dt, dt_total - t);
    const double inv_g_dt = 1.0 / (g * dt);
    for (int i = 0; i < N; ++i) Sx[i] = y[U::perm(i)];
    M::jac_full(y, k, dk, J, 1);
    for (int e = 0; e < N * N; ++e) {
      W[e] = 0.0;
      LU[e] = 0.0;
    }
    U::assemble_w(J, inv_g_dt, W);
    U::factorize(W, LU);
    for (int s = 0; s < S; ++s) {
      double rhs[N];
      const int base = s * (s - 1) / 2;
      if (s == 0 || Tab::NEWF[s]) {
        if (s == 0) {
          M::rates_full(y, k, fph);
        } else {
          for (int i = 0; i < N; ++i) {
            double v = Sx[i];
            for (int j = 0; j < s; ++j) {
              const double a = Tab::A[base + j];
              if (a == 0.0) continue;
              v = (a == 1.0) ? v + K[j][i] : v + a * K[j][i];
            }
            yph[U::perm(i)] = v;
          }
          M::rates_full(yph, k, fph);
        }
        for (int i = 0; i < N; ++i) F[i] = fph[U::perm(i)];
      }
      for (int i = 0; i < N; ++i) {
        double v = F[i];
        for (int j = 0; j < s; ++j) {
          const double c = Tab::C[base + j];
          if (c == 0.0) continue;
          v = v + (c / dt) * K[j][i];
        }
        rhs[i] = v;
      }
      U::solve(LU, rhs, K[s]);
    }
    // error norm over all N species in solver order (integrate.j2:265-322)
    double err_sq = 0.0;
    for (int i = 0; i < N; ++i) {
      double ynew = Sx[i];
      double yerr = 0.0;
      bool first = true;
      for (int j = 0; j < S; ++j) {
        const double m = Tab::M[j];
        if (m == 0.0) continue;
        ynew = ynew + ((m == 1.0) ? K[j][i] : m * K[j][i]);
      }
      for (int j = 0; j < S; ++j) {
        const double e = Tab::E[j];
        if (e == 0.0) continue;
        const double term = (e == 1.0) ? K[j][i] : e * K[j][i];
        yerr = first ? term : yerr + term;
        first = false;
      }
      const double ymax = fmax(fabs(Sx[i]), fabs(ynew));
      const double sci = atol[U::perm(i)] + rtol[U::perm(i)] * ymax;
      err_sq += (yerr / sci) * (yerr / sci);
    }
    double err = sqrt(err_sq / N);
    err = fmax(err, 1.0e-10);
    const double factor = step_factor<Tab>(err);
    if (err <= 1.0) {
      for (int i = 0; i < N; ++i) {
        double msum = 0.0;
        bool first = true;
        for (int j = 0; j < S; ++j) {
          const double m = Tab::M[j];
          if (m == 0.0) continue;
          const double term = (m == 1.0) ? K[j][i] : m * K[j][i];
          msum = first ? term : msum + term;
          first = false;
        }
        y[U::perm(i)] += msum;
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

// One cell per lane; inputs, status and counters as integrate_lane_cell.
template <class M, class U, class Tab>
__global__ void __launch_bounds__(64) integrate_unrolled_kernel(DeviceArgs a) {
  const long cell = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (cell >= a.ncells) return;
  static_assert(U::N == M::N, "unrolled and lane tables describe different species counts");
  double y[M::N], jv[M::NPHOTO], k[M::NK], dk[M::NK];
  for (int i = 0; i < M::N; ++i) y[i] = a.conc[cell + a.stride * i];
  for (int p = 0; p < M::NPHOTO; ++p)
    jv[p] = (a.jvals != nullptr && p < M::NUM_RUNTIME_PARAMETERS) ? a.jvals[cell * a.jcstride + a.jstride * p] : 0.0;
  const double T = a.temp ? a.temp[cell] : a.temp_default;
  const double rho = a.air_density ? a.air_density[cell] : a.rho_default;
  double Mtb = rho;
  for (int q = 0; q < M::NFIXED; ++q)
    if (M::fixed_phys(q) == M::THIRD_BODY) Mtb = y[M::fixed_phys(q)];
  if (a.status != nullptr && !all_finite(y, M::N)) {
    a.status[cell] = ERR_NON_FINITE_STATE;
    return;
  }
  M::rate_constants(T, Mtb, jv, a.cloud_water ? a.cloud_water[cell] : a.clw_default, k, dk);
  int status = SUCCESS, acc = 0, att = 0;
  for (int step = 0; step < a.steps && status == SUCCESS; ++step)
    status = integrate_cell_unrolled<M, U, Tab>(y, k, dk, a.dt, a.seed, a.atol, a.rtol, a.max_attempts, acc, att);
  bool fin = true;
  for (int v = 0; v < M::NV; ++v) {
    const double x = y[M::var_phys(v)];
    a.conc[cell + a.stride * M::var_phys(v)] = x;
    fin = fin && isfinite(x);
  }
  if (status == SUCCESS && !fin) status = ERR_NUMERICAL_FAILURE;
  if (a.status) a.status[cell] = status;
  if (a.accepted) a.accepted[cell] += acc;
  if (a.attempted) a.attempted[cell] += att;
}

template <class M, class U>
int launch_unrolled_integrate(const DeviceArgs& a, hipStream_t s) {
  if (a.ncells <= 0) return 0;
  const dim3 block(64);
  const dim3 grid((unsigned)((a.ncells + 63) / 64));
  switch (a.method) {
    case ROS2: integrate_unrolled_kernel<M, U, TabRos2><<<grid, block, 0, s>>>(a); break;
    case ROS3: integrate_unrolled_kernel<M, U, TabRos3><<<grid, block, 0, s>>>(a); break;
    case ROS4: integrate_unrolled_kernel<M, U, TabRos4><<<grid, block, 0, s>>>(a); break;
    case RODAS3: integrate_unrolled_kernel<M, U, TabRodas3><<<grid, block, 0, s>>>(a); break;
    case RODAS4: integrate_unrolled_kernel<M, U, TabRodas4><<<grid, block, 0, s>>>(a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace mkpp
