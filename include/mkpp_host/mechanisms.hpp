// mechanisms.hpp -- the compiled mechanisms' SolverKernels headers.
// Generated by mkpp_amd.gen.emit_host; do not edit.
//
// As in the reference (src/mkpp_host/include/mkpp_host/mechanisms.hpp:5-43),
// each mechanism's <name>.hpp is included when it is on the include path;
// MechanismRegistry then runs that mechanism with the header's
// generation-time solver options, and every other mechanism of
// libmkpp_amd.so with the library's defaults.
#pragma once
#include <string>

#include "dispatcher.hpp"

#if __has_include("chapman.hpp")
#include "chapman.hpp"
#define MKPP_HOST_HAS_CHAPMAN 1
#endif
#if __has_include("small_strato.hpp")
#include "small_strato.hpp"
#define MKPP_HOST_HAS_SMALL_STRATO 1
#endif
#if __has_include("carbon.hpp")
#include "carbon.hpp"
#define MKPP_HOST_HAS_CARBON 1
#endif
#if __has_include("saprc99.hpp")
#include "saprc99.hpp"
#define MKPP_HOST_HAS_SAPRC99 1
#endif
#if __has_include("saprcnov.hpp")
#include "saprcnov.hpp"
#define MKPP_HOST_HAS_SAPRCNOV 1
#endif
#if __has_include("decay.hpp")
#include "decay.hpp"
#define MKPP_HOST_HAS_DECAY 1
#endif
#if __has_include("decay_wave.hpp")
#include "decay_wave.hpp"
#define MKPP_HOST_HAS_DECAY_WAVE 1
#endif
#if __has_include("ts1.hpp")
#include "ts1.hpp"
#define MKPP_HOST_HAS_TS1 1
#endif
#if __has_include("gocart.hpp")
#include "gocart.hpp"
#define MKPP_HOST_HAS_GOCART 1
#endif

namespace mkpp::host::detail {

// The generation-time solver options of `name` if its header is included.
inline bool generated_options(const std::string& name, SolverOptions& o) {
#ifdef MKPP_HOST_HAS_CHAPMAN
  if (name == "chapman") {
    using K = ::mkpp::generated::chapman::SolverKernels<>;
    o.method = K::METHOD;
    o.atol = K::atol;
    o.rtol = K::rtol;
    o.initial_step_seed = K{}.initial_step_seed;
    return true;
  }
#endif
#ifdef MKPP_HOST_HAS_SMALL_STRATO
  if (name == "small_strato") {
    using K = ::mkpp::generated::small_strato::SolverKernels<>;
    o.method = K::METHOD;
    o.atol = K::atol;
    o.rtol = K::rtol;
    o.initial_step_seed = K{}.initial_step_seed;
    return true;
  }
#endif
#ifdef MKPP_HOST_HAS_CARBON
  if (name == "carbon") {
    using K = ::mkpp::generated::carbon::SolverKernels<>;
    o.method = K::METHOD;
    o.atol = K::atol;
    o.rtol = K::rtol;
    o.initial_step_seed = K{}.initial_step_seed;
    return true;
  }
#endif
#ifdef MKPP_HOST_HAS_SAPRC99
  if (name == "saprc99") {
    using K = ::mkpp::generated::saprc99::SolverKernels<>;
    o.method = K::METHOD;
    o.atol = K::atol;
    o.rtol = K::rtol;
    o.initial_step_seed = K{}.initial_step_seed;
    return true;
  }
#endif
#ifdef MKPP_HOST_HAS_SAPRCNOV
  if (name == "saprcnov") {
    using K = ::mkpp::generated::saprcnov::SolverKernels<>;
    o.method = K::METHOD;
    o.atol = K::atol;
    o.rtol = K::rtol;
    o.initial_step_seed = K{}.initial_step_seed;
    return true;
  }
#endif
#ifdef MKPP_HOST_HAS_DECAY
  if (name == "decay") {
    using K = ::mkpp::generated::decay::SolverKernels<>;
    o.method = K::METHOD;
    o.atol = K::atol;
    o.rtol = K::rtol;
    o.initial_step_seed = K{}.initial_step_seed;
    return true;
  }
#endif
#ifdef MKPP_HOST_HAS_DECAY_WAVE
  if (name == "decay_wave") {
    using K = ::mkpp::generated::decay_wave::SolverKernels<>;
    o.method = K::METHOD;
    o.atol = K::atol;
    o.rtol = K::rtol;
    o.initial_step_seed = K{}.initial_step_seed;
    return true;
  }
#endif
#ifdef MKPP_HOST_HAS_TS1
  if (name == "ts1") {
    using K = ::mkpp::generated::ts1::SolverKernels<>;
    o.method = K::METHOD;
    o.atol = K::atol;
    o.rtol = K::rtol;
    o.initial_step_seed = K{}.initial_step_seed;
    return true;
  }
#endif
#ifdef MKPP_HOST_HAS_GOCART
  if (name == "gocart") {
    using K = ::mkpp::generated::gocart::SolverKernels<>;
    o.method = K::METHOD;
    o.atol = K::atol;
    o.rtol = K::rtol;
    o.initial_step_seed = K{}.initial_step_seed;
    return true;
  }
#endif
  (void)name;
  (void)o;
  return false;
}

}  // namespace mkpp::host::detail
