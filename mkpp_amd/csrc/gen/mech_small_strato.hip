// Generated by mkpp_amd.gen.emit_hip: registers small_strato (one cell per lane;
// differentiation on the wave engine's tables of the same mechanism).
#include "small_strato.hpp"
#include "../runtime.hpp"
#include "../unrolled.hpp"
#include "../adjoint.hpp"
#include "small_strato_unrolled.hpp"
#include "small_strato_wcode.hpp"

MKPP_WAVE_TABLES(small_strato)
MKPP_LANE_MECHANISM(small_strato, 561.0, 673.0, 897.0, 825.0, 1445.0, (&launch_wave_diff<gen::small_strato_wave::Wave>))
