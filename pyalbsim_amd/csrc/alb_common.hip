// alb_common.hip — ABI bookkeeping: version, layout check, thread-local error.
#include <stdarg.h>
#include <stdio.h>

#include "alb_internal.h"

namespace alb {

static thread_local char g_error[512] = "";

int set_error(int status, const char* fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_error, sizeof g_error, fmt, ap);
  va_end(ap);
  return status;
}

int check_launch(const char* what) {
  const hipError_t e = hipGetLastError();
  if (e != hipSuccess) return set_error(ALB_ERR_HIP, "%s: launch failed: %s", what, hipGetErrorString(e));
  return ALB_OK;
}

}  // namespace alb

extern "C" int alb_abi_version(void) { return ALB_ABI_VERSION; }
extern "C" size_t alb_scene_size(void) { return sizeof(alb_scene_t); }
extern "C" const char* alb_last_error(void) { return alb::g_error; }
