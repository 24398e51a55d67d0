// Generated by mkpp_amd.gen.emit_hip: registers chapman (one cell per lane;
// differentiation on the wave engine's tables of the same mechanism).
#include "chapman.hpp"
#include "../runtime.hpp"
#include "../unrolled.hpp"
#include "../adjoint.hpp"
#include "chapman_unrolled.hpp"
#include "chapman_wcode.hpp"

MKPP_WAVE_TABLES(chapman)
MKPP_LANE_MECHANISM(chapman, 204.0, 249.0, 341.0, 305.0, 566.0, (&launch_wave_diff<gen::chapman_wave::Wave>))
