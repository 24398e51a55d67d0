// mkpp_host/mechanism_manager.hpp -- the reference's MechanismRegistry over
// the C ABI of libmkpp_amd.so (header-only).
//
// Replaces src/mkpp_host/include/mkpp_host/mechanism_manager.hpp:21-209:
// get_registered_mechanisms / has_mechanism / get_info / execute(name, state,
// params), with the same names, argument meaning and error behaviour (an
// unknown name throws std::runtime_error "FATAL ERROR: Unknown mechanism ...";
// mismatched per-cell extents throw std::invalid_argument).  Where the
// reference launches one Kokkos parallel_for per call and fences, execute()
// runs one GPU launch over every cell and returns when it has finished.
// A mechanism runs with its header's generation-time solver options when its
// <name>.hpp is on the include path (mkpp_host/mechanisms.hpp), else with the
// library's defaults.
#pragma once

#include <iostream>
#include <stdexcept>
#include <string>
#include <vector>

#include "dispatcher.hpp"
#include "mechanisms.hpp"

namespace mkpp {
namespace host {

struct MechanismInfo {
  std::string name;
  int num_species{0};
  int num_reactions{0};
  bool requires_env{false};  // rate constants depend on the per-cell environment
};

// Per-cell outcome of the last execute() of a mechanism over a cell count
// (no reference counterpart: the reference's execute reports no per-cell
// result; BatchErrorStatus is its batched driver's).
struct CellResults {
  std::vector<int> status, accepted, attempted;
};

class MechanismRegistry {
 public:
  static std::vector<std::string> get_registered_mechanisms() {
    std::vector<std::string> list;
    const int n = mkpp_amd_mechanism_count();
    for (int i = 0; i < n; ++i) list.emplace_back(mkpp_amd_mechanism_name(i));
    return list;
  }

  static bool has_mechanism(const std::string& name) { return mkpp_amd_has_mechanism(name.c_str()) != 0; }

  static MechanismInfo get_info(const std::string& name) {
    int ns = 0, nr = 0, nph = 0, np = 0, env = 0;
    if (mkpp_amd_mechanism_info(name.c_str(), &ns, &nr, &nph, &np, &env) != MKPP_SUCCESS) unknown(name);
    return {name, ns, nr, env != 0};
  }

  // execute(name, state, params): every cell of `state` (extent(0) cells x
  // species; device memory with cells contiguous, or host memory of any
  // layout when params are BasicHostExecutionParams<HostSpace>).
  template <class StateView, class MemorySpace>
  static void execute(const std::string& name, StateView state, const BasicHostExecutionParams<MemorySpace>& params) {
    if (!has_mechanism(name)) unknown(name);
    SolverOptions o;
    detail::generated_options(name, o);
    detail::execute_with(name, get_info(name).num_species, state, params, o);
  }

  // The per-cell status and step counts of the last execute() of `name` over
  // `num_cells` cells.
  static CellResults last_results(const std::string& name, std::size_t num_cells) {
    mkpp_handle_t h = detail::cached_handle(name, num_cells);
    if (!h) throw std::runtime_error("FATAL ERROR: no execute() of '" + name + "' over this many cells");
    CellResults r;
    r.status.resize(num_cells);
    r.accepted.resize(num_cells);
    r.attempted.resize(num_cells);
    const int rc = mkpp_amd_get_cell_status(h, r.status.data(), r.accepted.data(), r.attempted.data());
    if (rc != MKPP_SUCCESS) detail::fail(rc, h, "mkpp_amd_get_cell_status");
    return r;
  }

 private:
  [[noreturn]] static void unknown(const std::string& name) {
    std::string err = "FATAL ERROR: Unknown mechanism '" + name + "'. Supported: ";
    auto list = get_registered_mechanisms();
    for (std::size_t i = 0; i < list.size(); ++i) err += list[i] + (i + 1 < list.size() ? ", " : "");
    std::cerr << err << std::endl;
    throw std::runtime_error(err);
  }
};

}  // namespace host
}  // namespace mkpp
