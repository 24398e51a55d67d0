// Timing unit of the wave-engine autotuner (mkpp_amd/tune.py): one mechanism's
// integrate kernel for one tableau, compiled from a candidate's generated
// headers.  MECH (the mechanism's name) and TAB (a tableau type) are set on
// the compile line; the include path puts the candidate's headers first.  Not
// part of libmkpp_amd.so.
#include "wave_engine.hpp"

#define MKPP_TUNE_STR2(x) #x
#define MKPP_TUNE_STR(x) MKPP_TUNE_STR2(x)
#define MKPP_TUNE_CAT2(a, b) a##b
#define MKPP_TUNE_CAT(a, b) MKPP_TUNE_CAT2(a, b)
#include MKPP_TUNE_STR(MKPP_TUNE_CAT(MECH, _wcode.hpp))
#define MKPP_TUNE_TABLES(x) MKPP_WAVE_TABLES(x)
MKPP_TUNE_TABLES(MECH)

// The candidate's integrate kernel, with internal linkage: several candidates'
// libraries are loaded into one process, and a kernel with an external name
// (the product's wave_integrate_kernel) could bind every candidate's launch
// to the first library that defined it.
namespace {
template <class W, class Tab>
__global__ void __launch_bounds__(64) __attribute__((amdgpu_waves_per_eu(W::WAVES_PER_EU, W::WAVES_PER_EU)))
tune_kernel(mkpp::DeviceArgs a, long base) {
  extern __shared__ double lds[];
  const long cell = mkpp::wave_cell_of_block(base);
  if (cell < a.ncells) mkpp::wave_integrate_cell<W, Tab>(a, cell, lds, threadIdx.x);
}
}  // namespace

// The arguments of mkpp_amd_device_integrate for one tableau; 0 on success.
extern "C" int mkpp_tune_integrate(double* conc, long num_cells, long stride, const double* temp,
                                   const double* air_density, const double* rate_parameters, long parameter_stride,
                                   double dt, int steps, double initial_step_seed, const double* atol,
                                   const double* rtol, int* status, int* accepted, int* attempted, long max_attempts,
                                   void* stream) {
  using W = mkpp::gen::MKPP_TUNE_CAT(MECH, _wave)::Wave;
  if (!conc || !atol || !rtol || num_cells < 0 || stride < num_cells || steps < 1) return -103;
  mkpp::DeviceArgs a{};
  a.conc = conc;
  a.ncells = num_cells;
  a.stride = stride;
  a.temp = temp;
  a.air_density = air_density;
  a.jvals = rate_parameters;
  a.jstride = parameter_stride;
  a.dt = dt;
  a.steps = steps;
  a.seed = initial_step_seed;
  a.atol = atol;
  a.rtol = rtol;
  a.status = status;
  a.accepted = accepted;
  a.attempted = attempted;
  a.max_attempts = max_attempts;
  if (num_cells == 0) return 0;
  const size_t lds = sizeof(double) * W::LDS_DOUBLES;
  for (long base = 0; base < num_cells; base += mkpp::WAVE_CHUNK_CELLS) {
    const unsigned grid = mkpp::wave_chunk_grid(num_cells, base);
    mkpp::wave_allow_lds(&tune_kernel<W, mkpp::TAB>, lds);
    tune_kernel<W, mkpp::TAB><<<grid, 64, lds, (hipStream_t)stream>>>(a, base);
  }
  return hipGetLastError() == hipSuccess ? 0 : -104;
}
