// Generated by mkpp_amd.gen.emit_wave: registers saprc99 (one cell per wavefront).
#include "../adjoint.hpp"
#include "saprc99_wcode.hpp"

MKPP_WAVE_MECHANISM(saprc99, 36949.0, 40745.0, 47461.0, 46525.0, 63657.0)
