// Generated by mkpp_amd.gen.emit_hip: registers carbon (one cell per lane;
// differentiation on the wave engine's tables of the same mechanism).
#include "carbon.hpp"
#include "../runtime.hpp"
#include "../unrolled.hpp"
#include "../adjoint.hpp"
#include "carbon_unrolled.hpp"
#include "carbon_wcode.hpp"

MKPP_WAVE_TABLES(carbon)
MKPP_LANE_MECHANISM(carbon, 453.0, 606.0, 863.0, 731.0, 1459.0, (&launch_wave_diff<gen::carbon_wave::Wave>))
