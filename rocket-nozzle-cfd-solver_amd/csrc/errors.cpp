// errors.cpp — per-thread last-error text for the C ABI (nz_last_error).
#include <string>

#include "nz_internal.h"

namespace {
thread_local std::string g_last_error;
}

int nz_fail(int code, const std::string& msg) {
  g_last_error = msg;
  return code;
}

extern "C" const char* nz_last_error(void) { return g_last_error.c_str(); }
