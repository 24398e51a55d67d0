// Generated by mkpp_amd.gen.emit_hip: registers small_strato (one cell per lane).
#include "small_strato.hpp"
#include "../runtime.hpp"
#include "../unrolled.hpp"
#include "small_strato_unrolled.hpp"

MKPP_LANE_MECHANISM(small_strato, 561.0, 673.0, 897.0, 825.0, 1445.0)
