// Host-side mechanism registry: one MechOps record per compiled mechanism.
// Plays the role of the reference's MechanismRegistry
// (src/mkpp_host/include/mkpp_host/mechanism_manager.hpp:28-209), with HIP
// launches in place of Kokkos::parallel_for.
#pragma once
#include <hip/hip_runtime.h>

#include "mkpp_device.hpp"

namespace mkpp {

struct MechOps {
  const char* name;
  int num_species;
  int num_reactions;
  int num_photolysis;
  int num_runtime_parameters;
  int num_jacobian_entries;
  bool supports_runtime_environment;
  const char* const* species_names;
  int (*integrate)(const DeviceArgs& a, hipStream_t s);
  int (*eval_rates)(const EvalArgs& a, hipStream_t s);
  int (*eval_jacobian)(const EvalArgs& a, hipStream_t s);
  // per-cell accepted-step work estimate used by the batching scheduler:
  // algorithmic flops of one attempted step for each method (ros2..rodas4)
  double step_flops[5];
  // differentiation (wave engine only, null otherwise): mode 0 rate constants,
  // 1 checkpointed forward, 2 frozen-step replay, 3 adjoint
  int (*diff)(int mode, const DeviceArgs& a, const AdjArgs& g, hipStream_t s);
  int num_groups;    // flux groups (rate constants)
  int num_variable;  // variable species (checkpoint row length)
  // integrate_with_reduction (a.importance_threshold; tolerances by permuted position)
  int (*integrate_reduced)(const DeviceArgs& a, hipStream_t s);
  // validation backend: integrate over the dense unrolled LU (lane mechanisms
  // only, null otherwise)
  int (*integrate_unrolled)(const DeviceArgs& a, hipStream_t s);
};

const MechOps* find_mechanism(const char* name);
int mechanism_count();
const MechOps* mechanism_at(int i);

// Launch helper for lane-per-cell mechanisms.
template <class M>
int launch_lane_integrate(const DeviceArgs& a, hipStream_t s) {
  if (a.ncells <= 0) return 0;
  const dim3 block(128);
  const dim3 grid((unsigned)((a.ncells + 127) / 128));
  switch (a.method) {
    case ROS2: integrate_lane_kernel<M, TabRos2><<<grid, block, 0, s>>>(a); break;
    case ROS3: integrate_lane_kernel<M, TabRos3><<<grid, block, 0, s>>>(a); break;
    case ROS4: integrate_lane_kernel<M, TabRos4><<<grid, block, 0, s>>>(a); break;
    case RODAS3: integrate_lane_kernel<M, TabRodas3><<<grid, block, 0, s>>>(a); break;
    case RODAS4: integrate_lane_kernel<M, TabRodas4><<<grid, block, 0, s>>>(a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <class M>
int launch_lane_reduced(const DeviceArgs& a, hipStream_t s) {
  if (a.ncells <= 0) return 0;
  const dim3 block(128);
  const dim3 grid((unsigned)((a.ncells + 127) / 128));
  switch (a.method) {
    case ROS2: integrate_lane_reduced_kernel<M, TabRos2><<<grid, block, 0, s>>>(a); break;
    case ROS3: integrate_lane_reduced_kernel<M, TabRos3><<<grid, block, 0, s>>>(a); break;
    case ROS4: integrate_lane_reduced_kernel<M, TabRos4><<<grid, block, 0, s>>>(a); break;
    case RODAS3: integrate_lane_reduced_kernel<M, TabRodas3><<<grid, block, 0, s>>>(a); break;
    case RODAS4: integrate_lane_reduced_kernel<M, TabRodas4><<<grid, block, 0, s>>>(a); break;
    default: return -1;
  }
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <class M>
int launch_eval_rates(const EvalArgs& a, hipStream_t s) {
  if (a.ncells <= 0) return 0;
  eval_rates_kernel<M><<<dim3((unsigned)((a.ncells + 63) / 64)), dim3(64), 0, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

template <class M>
int launch_eval_jacobian(const EvalArgs& a, hipStream_t s) {
  if (a.ncells <= 0) return 0;
  eval_jacobian_kernel<M><<<dim3((unsigned)((a.ncells + 63) / 64)), dim3(64), 0, s>>>(a);
  return hipGetLastError() == hipSuccess ? 0 : -2;
}

}  // namespace mkpp

// Defines mkpp_ops_<name>() for a lane-per-cell mechanism whose generated
// headers declare mkpp::gen::<name>::Mech and mkpp::gen::<name>_unrolled::Unrolled
// (the unit includes unrolled.hpp).  DIFF: the differentiation entry (the wave
// engine's launch_wave_diff over the same mechanism's wave tables).
#define MKPP_LANE_MECHANISM(NAME, F0, F1, F2, F3, F4, DIFF)                                                  \
  namespace mkpp {                                                                                         \
  const MechOps* mkpp_ops_##NAME() {                                                                       \
    using M = gen::NAME::Mech;                                                                             \
    static const MechOps ops = {#NAME,                                                                     \
                                M::N,                                                                      \
                                M::NUM_REACTIONS,                                                          \
                                M::NUM_RUNTIME_PARAMETERS,                                                 \
                                M::NUM_RUNTIME_PARAMETERS,                                                 \
                                M::NJ,                                                                     \
                                M::SUPPORTS_RUNTIME_ENVIRONMENT,                                           \
                                M::SPECIES,                                                                \
                                &launch_lane_integrate<M>,                                                 \
                                &launch_eval_rates<M>,                                                     \
                                &launch_eval_jacobian<M>,                                                  \
                                {F0, F1, F2, F3, F4},                                                      \
                                DIFF,                                                                      \
                                M::NG,                                                                     \
                                M::NV,                                                                     \
                                &launch_lane_reduced<M>,                                                   \
                                &launch_unrolled_integrate<M, gen::NAME##_unrolled::Unrolled>};            \
    return &ops;                                                                                           \
  }                                                                                                        \
  }
