// mkpp_host/dispatcher.hpp -- the C++ host types of the reference's dispatcher
// over the C ABI of libmkpp_amd.so (header-only; no Kokkos, no HIP headers).
//
// Replaces src/mkpp_host/include/mkpp_host/dispatcher.hpp:15-98 for callers
// of MechanismRegistry::execute: the cell error codes, the execution
// parameters (BasicHostExecutionParams, with the same fields, defaults and
// per-cell override rules) and a minimal strided view.  Any view type with
// data(), extent(r) and stride(r) (a Kokkos::View included) can be passed
// where a view is taken.  The numerics run in libmkpp_amd.so.
#pragma once

#include <array>
#include <cstddef>
#include <mutex>
#include <stdexcept>
#include <string>
#include <tuple>
#include <type_traits>
#include <vector>

#include "../mkpp_amd.h"

namespace mkpp {
namespace host {

// dispatcher.hpp:15-21
enum CellErrorCode : int {
  SUCCESS = 0,
  ERR_SINGULAR_MATRIX = 1,
  ERR_NON_FINITE_STATE = 2,
  ERR_CONVERGENCE_FAILURE = 3,
  ERR_NUMERICAL_FAILURE = 4
};

// Where a view's data lives.  The default execution space of this build is
// the GPU, so the default memory space is device memory (as the reference's
// Kokkos::DefaultExecutionSpace::memory_space is on a GPU build).
struct HostSpace {};
struct DeviceSpace {};
using DefaultMemorySpace = DeviceSpace;

// A strided view of Rank dimensions over memory the caller owns.
template <class T, int Rank, class MemorySpace = DefaultMemorySpace>
class View {
 public:
  using value_type = T;
  using memory_space = MemorySpace;

  View() = default;
  // Rank 1: n0 contiguous elements.
  View(T* data, std::size_t n0) : data_(data), extent_{}, stride_{} {
    static_assert(Rank == 1, "this constructor makes a rank-1 view");
    extent_[0] = n0;
    stride_[0] = 1;
  }
  // Rank 2, LayoutLeft: element (i, j) at i + n0 * j.
  View(T* data, std::size_t n0, std::size_t n1) : data_(data), extent_{}, stride_{} {
    static_assert(Rank == 2, "this constructor makes a rank-2 view");
    extent_[0] = n0;
    extent_[1] = n1;
    stride_[0] = 1;
    stride_[1] = n0;
  }
  // Any extents and strides (LayoutStride).
  View(T* data, const std::array<std::size_t, Rank>& extents, const std::array<std::size_t, Rank>& strides)
      : data_(data), extent_(extents), stride_(strides) {}

  T* data() const { return data_; }
  std::size_t extent(int r) const { return r < Rank ? extent_[r] : 1; }
  std::size_t stride(int r) const { return r < Rank ? stride_[r] : 0; }
  static constexpr int rank() { return Rank; }
  T& operator()(std::size_t i) const { return data_[i * stride_[0]]; }
  T& operator()(std::size_t i, std::size_t j) const { return data_[i * stride_[0] + j * stride_[1]]; }

 private:
  T* data_ = nullptr;
  std::array<std::size_t, Rank> extent_{};
  std::array<std::size_t, Rank> stride_{};
};

// dispatcher.hpp:80-98: runtime execution parameters of a dispatch.  Per
// cell, T = cell_temperature(cell) when that view is non-empty, else temp;
// rh, air density and cloud water likewise; the forcing is the cell's row of
// cell_jvals when both its extents are non-zero, else jvals, else zeros.
template <class MemorySpace>
struct BasicHostExecutionParams {
  using memory_space = MemorySpace;

  double dt{60.0};                                  // integration interval per step, s
  View<const double, 1, MemorySpace> jvals{};       // shared photolysis / rate parameters
  double temp{288.15};                              // K
  double air_density{2.4476e19};                    // molec cm-3
  View<const double, 1, MemorySpace> cell_temperature;
  View<const double, 1, MemorySpace> cell_humidity;
  View<const double, 1, MemorySpace> cell_density;
  View<const double, 1, MemorySpace> cell_cloud_water;
  View<const double, 2, MemorySpace> cell_jvals;    // [cells x parameters], any strides
  double rh{0.5};
  double clw{0.0};
  int steps{1};                                     // integrate() calls per cell
  std::size_t jvals_extent{512};                    // size of the zero forcing when none is given
};

using HostExecutionParams = BasicHostExecutionParams<DefaultMemorySpace>;

// Solver options a dispatch runs with: a mechanism header's generation-time
// values, or the library's defaults (ROS-3, atol 1e-3, rtol 1e-6, seed 1e-6).
struct SolverOptions {
  int method = MKPP_METHOD_ROS3;
  const double* atol = nullptr;  // [num_species], physical order; nullptr: library default
  const double* rtol = nullptr;
  double initial_step_seed = 1.0e-6;
};

namespace detail {

inline std::string last_error(mkpp_handle_t h) {
  char buf[512] = {0};
  mkpp_amd_get_last_error(h, buf, sizeof buf);
  return buf;
}

[[noreturn]] inline void fail(int rc, mkpp_handle_t h, const std::string& what) {
  const std::string msg = what + ": " + (h ? last_error(h) : std::string()) + " (code " + std::to_string(rc) + ")";
  if (rc == MKPP_ERROR_INVALID_ARGUMENT) throw std::invalid_argument(msg);
  throw std::runtime_error(msg);
}

// One handle per (mechanism, cell count), kept for the process: a handle owns
// the device buffers of its cells' status and counters, so a dispatch loop
// reuses them instead of allocating per call.  Guarded by a mutex; a handle
// is used by one dispatch at a time.  The cache is never destroyed (its
// device buffers would otherwise be freed after the HIP runtime has shut
// down at exit); the process reclaims them.
struct HandleCache {
  std::mutex mutex;
  std::vector<std::tuple<std::string, std::size_t, mkpp_handle_t>> handles;
};

inline HandleCache& handle_cache() {
  static HandleCache* cache = new HandleCache();
  return *cache;
}

// The cached handle of the last dispatch of `name` over `cells` cells, or null.
inline mkpp_handle_t cached_handle(const std::string& name, std::size_t cells) {
  HandleCache& cache = handle_cache();
  std::lock_guard<std::mutex> lock(cache.mutex);
  for (auto& e : cache.handles)
    if (std::get<0>(e) == name && std::get<1>(e) == cells) return std::get<2>(e);
  return nullptr;
}

// Validation of the per-cell views (dispatcher.hpp:203-219).
template <class MemorySpace>
inline void check_extents(const BasicHostExecutionParams<MemorySpace>& p, std::size_t cells, int parameters) {
  for (const std::size_t e : {p.cell_temperature.extent(0), p.cell_humidity.extent(0), p.cell_density.extent(0),
                              p.cell_cloud_water.extent(0)})
    if (e != 0 && e != cells) throw std::invalid_argument("FATAL ERROR: per-cell meteorology extent does not match state");
  if (p.cell_jvals.extent(0) != 0 && p.cell_jvals.extent(0) != cells)
    throw std::invalid_argument("FATAL ERROR: per-cell forcing extent does not match state");
  if (p.cell_jvals.extent(0) != 0 && p.cell_jvals.extent(1) < static_cast<std::size_t>(parameters))
    throw std::invalid_argument("FATAL ERROR: per-cell forcing has too few runtime parameters");
}

// Runs every cell of `state` (extent(0) cells x num_species) with options `o`:
// device memory goes straight to mkpp_amd_execute; host memory is gathered
// into the C ABI's LayoutLeft buffers and integrated through the handle API.
template <class StateView, class MemorySpace>
void execute_with(const std::string& name, int num_species, StateView state,
                  const BasicHostExecutionParams<MemorySpace>& params, const SolverOptions& o) {
  const std::size_t cells = state.extent(0);
  if (cells == 0) return;
  if (state.extent(1) != static_cast<std::size_t>(num_species))
    throw std::invalid_argument("FATAL ERROR: state extent(1) does not match the mechanism's species count");
  for (int r = 2; r < StateView::rank(); ++r)
    if (state.extent(r) != 1) throw std::invalid_argument("FATAL ERROR: state extents beyond the species must be 1");
  int ns = 0, nr = 0, nph = 0, np = 0, env = 0;
  if (mkpp_amd_mechanism_info(name.c_str(), &ns, &nr, &nph, &np, &env) != MKPP_SUCCESS || ns != num_species)
    throw std::runtime_error("FATAL ERROR: mechanism '" + name + "' is not in libmkpp_amd.so");
  check_extents(params, cells, np);

  HandleCache& cache = handle_cache();
  std::lock_guard<std::mutex> lock(cache.mutex);
  mkpp_handle_t h = nullptr;
  for (auto& e : cache.handles)
    if (std::get<0>(e) == name && std::get<1>(e) == cells) h = std::get<2>(e);
  if (!h) {
    int rc = mkpp_amd_create_handle(name.c_str(), &h, cells);
    if (rc != MKPP_SUCCESS) fail(rc, nullptr, "mkpp_amd_create_handle");
    if ((rc = mkpp_amd_initialize(h)) != MKPP_SUCCESS) fail(rc, h, "mkpp_amd_initialize");
    cache.handles.emplace_back(name, cells, h);
  }
  std::vector<double> at, rt;
  if (!o.atol || !o.rtol) {
    at.assign(num_species, 1.0e-3);
    rt.assign(num_species, 1.0e-6);
  }
  int rc = mkpp_amd_set_solver_options(h, o.method, o.atol ? o.atol : at.data(), o.rtol ? o.rtol : rt.data(),
                                       o.initial_step_seed);
  if (rc != MKPP_SUCCESS) fail(rc, h, "mkpp_amd_set_solver_options");

  auto ptr = [](const auto& v) -> const double* { return v.extent(0) ? v.data() : nullptr; };
  if constexpr (std::is_same_v<MemorySpace, DeviceSpace>) {
    if (state.stride(0) != 1)
      throw std::invalid_argument("FATAL ERROR: device state must hold cells contiguously (stride(0) == 1)");
    mkpp_amd_execution_params p{};
    p.dt = params.dt;
    p.steps = params.steps;
    p.jvals = ptr(params.jvals);
    p.jvals_extent = static_cast<long>(params.jvals.extent(0));
    p.temp = params.temp;
    p.air_density = params.air_density;
    p.rh = params.rh;
    p.clw = params.clw;
    p.cell_temperature = ptr(params.cell_temperature);
    p.cell_humidity = ptr(params.cell_humidity);
    p.cell_density = ptr(params.cell_density);
    p.cell_cloud_water = ptr(params.cell_cloud_water);
    p.cell_jvals = params.cell_jvals.extent(0) ? params.cell_jvals.data() : nullptr;
    p.cell_jvals_extent0 = static_cast<long>(params.cell_jvals.extent(0));
    p.cell_jvals_extent1 = static_cast<long>(params.cell_jvals.extent(1));
    p.cell_jvals_stride0 = static_cast<long>(params.cell_jvals.stride(0));
    p.cell_jvals_stride1 = static_cast<long>(params.cell_jvals.stride(1));
    rc = mkpp_amd_execute(h, const_cast<double*>(state.data()), static_cast<long>(cells),
                          static_cast<long>(state.stride(1)), &p, nullptr);
    if (rc != MKPP_SUCCESS) fail(rc, h, "mkpp_amd_execute");
  } else {
    // host memory: LayoutLeft staging buffers with the per-cell rules applied
    std::vector<double> conc(cells * num_species), T(cells), rho(cells), pres(cells, 101325.0),
        jv(cells * static_cast<std::size_t>(np > 0 ? np : 1), 0.0);
    for (std::size_t c = 0; c < cells; ++c)
      for (int s = 0; s < num_species; ++s) conc[c + cells * s] = state.data()[c * state.stride(0) + s * state.stride(1)];
    for (std::size_t c = 0; c < cells; ++c) {
      T[c] = params.cell_temperature.extent(0) ? params.cell_temperature(c) : params.temp;
      rho[c] = params.cell_density.extent(0) ? params.cell_density(c) : params.air_density;
      for (int q = 0; q < np; ++q) {
        double v = 0.0;
        if (params.cell_jvals.extent(0) && params.cell_jvals.extent(1)) v = params.cell_jvals(c, q);
        else if (params.jvals.extent(0) > static_cast<std::size_t>(q)) v = params.jvals(q);
        jv[c + cells * q] = v;
      }
    }
    if ((rc = mkpp_amd_set_state_ptrs(h, conc.data(), T.data(), pres.data(), rho.data())) != MKPP_SUCCESS)
      fail(rc, h, "mkpp_amd_set_state_ptrs");
    if (np > 0 && (rc = mkpp_amd_set_rate_parameter_ptrs(h, jv.data(), static_cast<std::size_t>(np))) != MKPP_SUCCESS)
      fail(rc, h, "mkpp_amd_set_rate_parameter_ptrs");
    if ((rc = mkpp_amd_integrate_steps(h, params.dt, params.steps)) != MKPP_SUCCESS) fail(rc, h, "mkpp_amd_integrate");
    for (std::size_t c = 0; c < cells; ++c)
      for (int s = 0; s < num_species; ++s)
        const_cast<double*>(state.data())[c * state.stride(0) + s * state.stride(1)] = conc[c + cells * s];
  }
}

}  // namespace detail
}  // namespace host
}  // namespace mkpp
