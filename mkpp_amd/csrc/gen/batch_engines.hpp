// Generated by mkpp_amd.gen.emit_batch: every bundled mechanism in one persistent kernel.
#pragma once
#include "../batch.hpp"
#include "chapman.hpp"
#include "small_strato.hpp"
#include "carbon.hpp"
#include "saprc99_wcode.hpp"
#include "saprcnov_wcode.hpp"
MKPP_WAVE_TABLES(saprc99)
MKPP_WAVE_TABLES(saprcnov)

namespace mkpp {
struct BatchEngines {
  template <class Tab>
  static __device__ __forceinline__ void run(int engine, const DeviceArgs& a, long item, int m, double* lds) {
    // opaque per item: nothing lane-dependent is hoisted out of the queue loop
    // (it would stay live in registers across every engine's code)
    int lane = threadIdx.x;
    asm volatile("" : "+v"(lane));
    switch (engine) {
      case 0: batch_lane_item<gen::chapman::Mech, Tab>(a, item, m); break;
      case 1: batch_lane_item<gen::small_strato::Mech, Tab>(a, item, m); break;
      case 2: batch_lane_item<gen::carbon::Mech, Tab>(a, item, m); break;
      case 3: wave_integrate_cell<gen::saprc99_wave::Wave, Tab>(a, item, lds, lane); break;
      case 4: wave_integrate_cell<gen::saprcnov_wave::Wave, Tab>(a, item, lds, lane); break;
      default: break;
    }
  }
};
}  // namespace mkpp
