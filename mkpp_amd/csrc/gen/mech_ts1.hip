// Generated by mkpp_amd.gen.emit_wave: registers ts1 (one cell per wavefront).
#include "../adjoint.hpp"
#include "ts1_wcode.hpp"

MKPP_WAVE_MECHANISM_NODIFF(ts1, 143700.0, 155095.0, 173126.0, 170618.0, 215406.0)
