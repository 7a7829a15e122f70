// C-ABI plumbing: thread-local error text, launch checks, version query.
#include "common.h"

#include <cstring>

namespace mom {

static thread_local std::string g_error;

void set_error(const std::string& msg) { g_error = msg; }

int fail(const std::string& msg) {
  g_error = msg;
  return -1;
}

int check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return fail(std::string(what) + ": " + hipGetErrorString(e));
  return 0;
}

}  // namespace mom

extern "C" const char* mom_last_error(void) { return mom::g_error.c_str(); }

extern "C" int mom_abi_version(void) { return 1; }
