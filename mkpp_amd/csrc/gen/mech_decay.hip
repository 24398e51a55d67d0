// Generated by mkpp_amd.gen.emit_hip: registers decay (one cell per lane;
// differentiation on the wave engine's tables of the same mechanism).
#include "decay.hpp"
#include "../runtime.hpp"
#include "../unrolled.hpp"
#include "../adjoint.hpp"
#include "decay_unrolled.hpp"
#include "decay_wcode.hpp"

MKPP_WAVE_TABLES(decay)
MKPP_LANE_MECHANISM(decay, 30.0, 41.0, 61.0, 49.0, 108.0, (&launch_wave_diff<gen::decay_wave::Wave>))
