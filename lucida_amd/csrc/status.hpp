// Status plumbing shared by the C-ABI entry points: a thread-local message for
// the last failure, standing in for the String a GpuError carries
// (crates/gpu/src/lib.rs:100-112).
#pragma once
#include "../../include/lucida.h"

namespace lucida {

lucida_status set_error(lucida_status code, const char* msg);
void clear_error();

}  // namespace lucida
