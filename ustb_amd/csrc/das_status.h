// das_status.h — filling the C-ABI's ustbx_status (internal to the library).
#pragma once
#include <hip/hip_runtime.h>
#include <stdarg.h>
#include <stdio.h>

#include "../../include/ustbx_das.h"

namespace {

inline int set_status(ustbx_status* st, int code, const char* fmt, ...) {
  if (st) {
    st->code = code;
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(st->message, sizeof(st->message), fmt, ap);
    va_end(ap);
  }
  return code;
}

inline int ok(ustbx_status* st) {
  if (st) {
    st->code = USTBX_OK;
    st->message[0] = 0;
  }
  return USTBX_OK;
}

inline int hip_fail(ustbx_status* st, hipError_t e, const char* what) {
  return set_status(st, USTBX_E_GPU, "HIP error: %s in %s", hipGetErrorString(e), what);
}

}  // namespace
