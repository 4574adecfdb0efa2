// Error reporting and small runtime queries of the nbx C ABI.
#include <hip/hip_runtime.h>
#include <cstdarg>
#include <cstdio>
#include "../../include/nbx.h"

namespace nbx {

static thread_local char g_err[1024] = "";

void set_error(const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
}

int fail(int code, const char* fmt, ...)
{
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(g_err, sizeof(g_err), fmt, ap);
  va_end(ap);
  return code;
}

}  // namespace nbx

extern "C" const char* nbx_last_error(void) { return nbx::g_err; }

extern "C" int nbx_version(void) { return 100; }

extern "C" int nbx_device_synchronize(void)
{
  const hipError_t e = hipDeviceSynchronize();
  if (e != hipSuccess) return nbx::fail(NBX_ERR_HIP, "hipDeviceSynchronize: %s", hipGetErrorString(e));
  return NBX_SUCCESS;
}
