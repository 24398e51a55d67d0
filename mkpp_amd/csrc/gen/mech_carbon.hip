// Generated by mkpp_amd.gen.emit_hip: registers carbon (one cell per lane).
#include "carbon.hpp"
#include "../runtime.hpp"
#include "../unrolled.hpp"
#include "carbon_unrolled.hpp"

MKPP_LANE_MECHANISM(carbon, 453.0, 606.0, 863.0, 731.0, 1459.0)
