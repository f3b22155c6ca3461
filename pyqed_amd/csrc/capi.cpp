// C ABI bookkeeping shared by every entry point: thread-local error text,
// version and device count.
#include <hip/hip_runtime.h>

#include <string>

#include "common.h"

namespace pqa {

static thread_local std::string g_last_error = "no error";

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

}  // namespace pqa

extern "C" const char* pyqed_last_error(void) { return pqa::g_last_error.c_str(); }

extern "C" int pyqed_version(void) { return 100; }

extern "C" int pyqed_device_count(int* count) {
  if (!count) return pqa::fail(PYQED_EINVAL, "device_count: null pointer");
  int n = 0;
  if (hipGetDeviceCount(&n) != hipSuccess) n = 0;
  *count = n;
  return PYQED_OK;
}
