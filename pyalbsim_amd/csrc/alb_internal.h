// alb_internal.h — host-side helpers shared by the library's translation units.
#pragma once

#include <hip/hip_runtime.h>

#include "../../include/alb_abi.h"

namespace alb {

// Record an error message for alb_last_error() and return the status.
int set_error(int status, const char* fmt, ...);

// Status of the last kernel launch on this thread.
int check_launch(const char* what);

}  // namespace alb
