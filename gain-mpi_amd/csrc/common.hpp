// Shared host-side helpers of the gain_engine library: error reporting and
// HIP call checking.  Every C-ABI entry point converts exceptions thrown
// here into a gain_status plus a thread-local message.
#pragma once
#include <hip/hip_runtime.h>

#include <cstdarg>
#include <cstdio>
#include <stdexcept>
#include <string>

#include "gain_engine.h"

namespace gain {

struct Error : std::runtime_error {
  int code;
  Error(int c, const std::string& m) : std::runtime_error(m), code(c) {}
};

[[noreturn]] inline void fail(int code, const char* fmt, ...) {
  char buf[512];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  throw Error(code, buf);
}

#define GAIN_HIP(call)                                                    \
  do {                                                                    \
    hipError_t e_ = (call);                                               \
    if (e_ != hipSuccess)                                                 \
      ::gain::fail(GAIN_ERR_HIP, "%s failed: %s (%s:%d)", #call,          \
                   hipGetErrorString(e_), __FILE__, __LINE__);            \
  } while (0)

#define GAIN_CHECK_ARG(cond, ...)                                         \
  do {                                                                    \
    if (!(cond)) ::gain::fail(GAIN_ERR_ARG, __VA_ARGS__);                 \
  } while (0)

void set_last_error(const char* msg);

// Wraps a C-ABI body: returns GAIN_OK or the error code, recording the
// message for gain_last_error().
template <class F>
int guarded(F&& f) {
  try {
    f();
    set_last_error("");
    return GAIN_OK;
  } catch (const Error& e) {
    set_last_error(e.what());
    return e.code;
  } catch (const std::bad_alloc&) {
    set_last_error("host allocation failed");
    return GAIN_ERR_NOMEM;
  } catch (const std::exception& e) {
    set_last_error(e.what());
    return GAIN_ERR_ARG;
  }
}

// Launch check: surfaces configuration errors of the launch just made.
inline void check_launch(const char* what) {
  hipError_t e = hipGetLastError();
  if (e != hipSuccess)
    fail(GAIN_ERR_HIP, "launch of %s failed: %s", what, hipGetErrorString(e));
}

}  // namespace gain
