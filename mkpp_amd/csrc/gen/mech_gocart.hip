// Generated by mkpp_amd.gen.emit_hip: registers gocart (one cell per lane;
// no differentiation kernels).
#include "gocart.hpp"
#include "../runtime.hpp"
#include "../unrolled.hpp"
#include "gocart_unrolled.hpp"

MKPP_LANE_MECHANISM(gocart, 2902.0, 3490.0, 4473.0, 4161.0, 6782.0, nullptr)
