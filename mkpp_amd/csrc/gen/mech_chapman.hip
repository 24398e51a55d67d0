// Generated by mkpp_amd.gen.emit_hip: registers chapman (one cell per lane).
#include "chapman.hpp"
#include "../runtime.hpp"
#include "../unrolled.hpp"
#include "chapman_unrolled.hpp"

MKPP_LANE_MECHANISM(chapman, 204.0, 249.0, 341.0, 305.0, 566.0)
