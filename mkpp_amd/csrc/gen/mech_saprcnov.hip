// Generated by mkpp_amd.gen.emit_wave: registers saprcnov (one cell per wavefront).
#include "../adjoint.hpp"
#include "saprcnov_wcode.hpp"

MKPP_WAVE_MECHANISM(saprcnov, 29302.0, 32871.0, 39131.0, 38015.0, 54156.0)
