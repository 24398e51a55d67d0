// Generated by mkpp_amd.gen.emit_hip: registers decay (one cell per lane).
#include "decay.hpp"
#include "../runtime.hpp"
#include "../unrolled.hpp"
#include "decay_unrolled.hpp"

MKPP_LANE_MECHANISM(decay, 30.0, 41.0, 61.0, 49.0, 108.0)
