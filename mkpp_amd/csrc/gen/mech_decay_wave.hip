// Generated by mkpp_amd.gen.emit_wave: registers decay_wave (one cell per wavefront).
#include "../adjoint.hpp"
#include "decay_wave_wcode.hpp"

MKPP_WAVE_MECHANISM(decay_wave, 30.0, 41.0, 61.0, 49.0, 108.0)
