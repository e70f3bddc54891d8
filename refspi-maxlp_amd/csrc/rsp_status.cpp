// Thread-local error message and status helpers of the C-ABI.
#include <string>

#include "rsp_common.hpp"
#include "refspi_amd.h"

namespace rsp {

static thread_local std::string g_last_error;

void set_error(const std::string& msg) { g_last_error = msg; }

int fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  g_last_error = std::string(what) + ": " + hipGetErrorString(e);
  return (int)e;
}

}  // namespace rsp

extern "C" const char* rsp_last_error(void) { return rsp::g_last_error.c_str(); }

extern "C" int rsp_version(void) { return 100; }
