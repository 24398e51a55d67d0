// Generated by mkpp_amd.gen.emit_batch: the persistent batch kernel for Ros3.
#include "batch_engines.hpp"

MKPP_BATCH_METHOD(TabRos3)
